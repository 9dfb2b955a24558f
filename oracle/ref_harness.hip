// Launchers of the reference's own kernels (compiled from kernel.cl by
// oracle/build_ref.sh) -- TEST INFRASTRUCTURE.
//
// ref_kdtree_intersect reproduces the first intersect of TraceRunner::run
// (trace_runner.cpp:192-206): every ray is a 144-byte unit_data record
// (types.h.cl:8-17) with orig_id = index, strength = 1, start_p and in_dir
// set; v_s0 = 0..n-1; sample_n = 1 and a black background; 256-wide groups
// over (n/256 + 1)*256 items.  Returns the geometry int4 and
// intersect_number the kernel writes into each record (left at -1 / -42 on
// a miss).
//
// ref_render reproduces the whole of TraceRunner::run up to results_mem
// (trace_runner.cpp:116-327) with the reference's kernels and its host
// protocol for the list sizes.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

namespace {
struct F3 {
    float x, y, z, w;
};
struct Unit {  // unit_data, 144 bytes (offsets 0,16,32,48,64,80,96,112,128,132)
    int32_t orig_id;
    int32_t pad0[3];
    F3 strength, start_p, in_dir;
    int32_t geometry[4];
    float intersect_number;
    int32_t pad1[3];
    F3 intersect_p, normal;
    float optical_density, shininess;
    int32_t pad2[2];
};
static_assert(sizeof(Unit) == 144, "unit_data layout");

thread_local std::string g_err;

bool ok(hipError_t e, const char *what) {
    if (e == hipSuccess) return true;
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return false;
}

template <class T>
T *dev_copy(const T *src, size_t n, std::vector<void *> &keep) {
    void *p = nullptr;
    if (!ok(hipMalloc(&p, sizeof(T) * (n ? n : 1)), "hipMalloc")) return nullptr;
    keep.push_back(p);
    if (n && !ok(hipMemcpy(p, src, sizeof(T) * n, hipMemcpyHostToDevice), "upload")) return nullptr;
    return static_cast<T *>(p);
}
}  // namespace

extern "C" {

const char *ref_last_error(void) { return g_err.c_str(); }

// points: 4 floats per point; mesh: int4; nodes: 48-byte KDTreeNodeHeader;
// org/dir: xyz triples.  geometry: 4 ints per ray, t: 1 float per ray.
// kernel_ms (may be NULL): device time of the kdtree_intersect launch, from
// HIP events around it.
int ref_kdtree_intersect_timed(const char *co_path, const float *points, int64_t n_points,
                               const int32_t *mesh, int64_t n_tris, const void *nodes,
                               int64_t n_nodes, const int32_t *leaf, int64_t n_leaf,
                               const float *org, const float *dir, int32_t n, int32_t *geometry,
                               float *t, float *kernel_ms) {
    std::vector<void *> keep;
    hipModule_t mod = nullptr;
    int rc = -1;
    do {
        if (!ok(hipModuleLoad(&mod, co_path), "hipModuleLoad")) break;
        hipFunction_t fn;
        if (!ok(hipModuleGetFunction(&fn, mod, "kdtree_intersect"), "hipModuleGetFunction")) break;
        std::vector<Unit> data(n);
        std::vector<int32_t> s0(n);
        for (int32_t i = 0; i < n; ++i) {
            Unit &u = data[i];
            std::memset(&u, 0, sizeof(u));
            u.orig_id = i;
            u.strength = F3{1, 1, 1, 0};
            u.start_p = F3{org[3 * i], org[3 * i + 1], org[3 * i + 2], 0};
            u.in_dir = F3{dir[3 * i], dir[3 * i + 1], dir[3 * i + 2], 0};
            for (int k = 0; k < 4; ++k) u.geometry[k] = -1;
            u.intersect_number = -42.f;
            s0[i] = i;
        }
        int32_t sizes[10] = {n, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        std::vector<float> zeros(3 * (size_t)n, 0.f);
        int32_t *d_sizes = dev_copy(sizes, 10, keep);
        Unit *d_data = dev_copy(data.data(), data.size(), keep);
        int32_t *d_s0 = dev_copy(s0.data(), s0.size(), keep);
        int32_t *d_s1 = dev_copy(s0.data(), s0.size(), keep);
        const float *d_points = dev_copy(points, 4 * (size_t)n_points, keep);
        const int32_t *d_mesh = dev_copy(mesh, 4 * (size_t)n_tris, keep);
        const int32_t *d_leaf = dev_copy(leaf, (size_t)n_leaf, keep);
        const uint8_t *d_nodes = dev_copy(static_cast<const uint8_t *>(nodes), 48 * (size_t)n_nodes, keep);
        float *d_result = dev_copy(zeros.data(), zeros.size(), keep);
        if (!d_sizes || !d_data || !d_s0 || !d_s1 || !d_points || !d_mesh || !d_leaf || !d_nodes ||
            !d_result)
            break;
        int32_t kd_node_size = (int32_t)n_nodes;
        F3 bg{0, 0, 0, 0};
        int32_t sample_n = 1;
        void *args[] = {&d_sizes, &d_data, &d_s0, &d_s1, &d_points, &d_mesh, &d_leaf,
                        &d_nodes, &kd_node_size, &d_result, &bg, &sample_n};
        const unsigned groups = (unsigned)(n / 256 + 1);
        hipEvent_t ev[2] = {nullptr, nullptr};
        if (kernel_ms && (!ok(hipEventCreate(&ev[0]), "hipEventCreate") ||
                          !ok(hipEventCreate(&ev[1]), "hipEventCreate")))
            break;
        if (kernel_ms && !ok(hipEventRecord(ev[0], nullptr), "hipEventRecord")) break;
        if (!ok(hipModuleLaunchKernel(fn, groups, 1, 1, 256, 1, 1, 0, nullptr, args, nullptr),
                "hipModuleLaunchKernel"))
            break;
        if (kernel_ms && !ok(hipEventRecord(ev[1], nullptr), "hipEventRecord")) break;
        if (!ok(hipDeviceSynchronize(), "kdtree_intersect")) break;
        if (kernel_ms) {
            if (!ok(hipEventElapsedTime(kernel_ms, ev[0], ev[1]), "hipEventElapsedTime")) break;
            (void)hipEventDestroy(ev[0]);
            (void)hipEventDestroy(ev[1]);
        }
        if (!ok(hipMemcpy(data.data(), d_data, sizeof(Unit) * n, hipMemcpyDeviceToHost), "download"))
            break;
        for (int32_t i = 0; i < n; ++i) {
            std::memcpy(geometry + 4 * (size_t)i, data[i].geometry, 16);
            t[i] = data[i].intersect_number;
        }
        rc = 0;
    } while (false);
    for (void *p : keep) (void)hipFree(p);
    if (mod) (void)hipModuleUnload(mod);
    return rc;
}

// TraceRunner::run (trace_runner.cpp:116-327) up to results_mem, with the
// reference's kernels: set_viewdirs[_vr] over (W*ssx, H*ssy), the first
// kdtree_intersect with sample_n = sample whose S1 list is kept, then per
// sample the restore of the records and of S1 and `depth` levels of
// kdtree_intersect (skipped while S0 is empty), s1_run and the four s2
// kernels, the host reading v_sizes after each stage and zeroing the
// consumed counts exactly as run() does.  Differences, none of which the
// kernels can see: results_mem starts at zero (run() leaves it
// uninitialised) and the per-slot seeds are the caller's instead of
// rand().  co_path is the build whose s1_run takes its texture argument as
// an int (ref_kernels_notex.co: an OpenCL image object cannot be passed
// through HIP), so every material must be untextured (checked; -2
// otherwise).
//
// Arrays as hpt_scene_desc holds them: points and normals 4 floats each,
// texcoords 2, mesh and lights int4, materials 96 bytes, nodes 48 bytes.
// seeds: W*ssx*H*ssy values; results: 3 floats per camera ray; s0_sizes
// (may be NULL): sample*depth ints, the S0 count each level's intersect
// ran over; loop_seconds (may be NULL): host wall time from set_viewdirs to
// the end of the last level (the part of run() after the constructor's
// uploads, without get_image); primary_geometry (may be NULL): 4 ints per
// camera ray, the int4 of the triangle its primary intersect found
// (scene_mesh[geo_id], kernel.cl:187-192), or -1s for a miss.
int ref_render(const char *co_path, const float *points, int64_t n_points, const float *normals,
               int64_t n_normals, const float *texcoords, int64_t n_texcoords,
               const int32_t *mesh, int64_t n_tris, const void *materials, int64_t n_materials,
               const int32_t *lights, int64_t n_lights, const void *nodes, int64_t n_nodes,
               const int32_t *leaf, int64_t n_leaf, const float *view, const float *up,
               const float *right, float width, float height, int32_t sample_x,
               int32_t sample_y, int32_t ssx, int32_t ssy, const float *background,
               int32_t sample, int32_t depth, int32_t disable_diffuse, const int64_t *seeds,
               float *results, int32_t *s0_sizes, double *loop_seconds,
               int32_t *primary_geometry) {
    std::vector<void *> keep;
    hipModule_t mod = nullptr;
    int rc = -1;
    do {
        const uint8_t *mats = static_cast<const uint8_t *>(materials);
        bool textured = false;
        for (int64_t m = 0; m < n_materials; ++m) {
            int32_t diffuse_tex, ambient_tex;  // Material offsets 76 and 80 (types.h.cl)
            std::memcpy(&diffuse_tex, mats + 96 * m + 76, 4);
            std::memcpy(&ambient_tex, mats + 96 * m + 80, 4);
            textured = textured || diffuse_tex >= 0 || ambient_tex >= 0;
        }
        if (textured) {
            g_err = "ref_render: textured materials are not supported";
            rc = -2;
            break;
        }
        const int32_t gx = sample_x * ssx, gy = sample_y * ssy;
        const int32_t R = gx * gy;
        if (sample_x <= 0 || sample_y <= 0 || ssx <= 0 || ssy <= 0 || sample <= 0 || depth <= 0) {
            g_err = "ref_render: bad view";
            break;
        }
        // set_viewdirs_vr indexes its records with sample_x * ssy as the row
        // length over a range of sample_x * ssx columns: in bounds only when
        // the two supersampling factors are equal.
        if (height < 0 && ssx != ssy) {
            g_err = "ref_render: the equirectangular camera needs ssx == ssy";
            break;
        }
        if (!ok(hipModuleLoad(&mod, co_path), "hipModuleLoad")) break;
        const char *names[] = {"set_viewdirs", "set_viewdirs_vr", "kdtree_intersect", "s1_run",
                               "s2_refract_run", "s2_specular_run", "s2_diffuse_run",
                               "s2_light_run"};
        hipFunction_t fn[8];
        bool got = true;
        for (int k = 0; k < 8 && got; ++k)
            got = ok(hipModuleGetFunction(&fn[k], mod, names[k]), names[k]);
        if (!got) break;

        auto alloc = [&](size_t bytes) -> void * {
            void *p = nullptr;
            if (!ok(hipMalloc(&p, bytes < 64 ? 64 : bytes), "hipMalloc")) return nullptr;
            keep.push_back(p);
            if (!ok(hipMemset(p, 0, bytes < 64 ? 64 : bytes), "hipMemset")) return nullptr;
            return p;
        };
        auto upload = [&](const void *src, size_t bytes) -> void * {
            void *p = alloc(bytes);
            if (p && bytes && !ok(hipMemcpy(p, src, bytes, hipMemcpyHostToDevice), "upload"))
                return nullptr;
            return p;
        };
        void *d_points = upload(points, 16 * (size_t)n_points);
        void *d_normals = upload(normals, 16 * (size_t)n_normals);
        void *d_texcoords = upload(texcoords, 8 * (size_t)n_texcoords);
        void *d_mesh = upload(mesh, 16 * (size_t)n_tris);
        void *d_mats = upload(materials, 96 * (size_t)n_materials);
        void *d_lights = upload(lights, 16 * (size_t)n_lights);
        void *d_nodes = upload(nodes, 48 * (size_t)n_nodes);
        void *d_leaf = upload(leaf, 4 * (size_t)n_leaf);
        int32_t no_texture = 0;  // s1_run's texture argument (an int in this build)
        void *d_initial = alloc(sizeof(Unit) * (size_t)R);
        void *d_data = alloc(sizeof(Unit) * (size_t)R);
        void *d_results = alloc(12 * (size_t)R);
        void *d_sizes = alloc(40);
        void *d_seeds = upload(seeds, 8 * (size_t)R);
        std::vector<int32_t> iota(R);
        for (int32_t i = 0; i < R; ++i) iota[i] = i;
        void *d_cache = upload(iota.data(), 4 * (size_t)R);
        void *d_stage[6];
        bool all = d_points && d_normals && d_texcoords && d_mesh && d_mats && d_lights &&
                   d_nodes && d_leaf && d_initial && d_data && d_results && d_sizes &&
                   d_seeds && d_cache;
        for (int k = 0; k < 6; ++k) {
            d_stage[k] = alloc(4 * (size_t)R);
            all = all && d_stage[k];
        }
        if (!all) break;

        if (!ok(hipDeviceSynchronize(), "uploads")) break;
        const auto t_loop = std::chrono::steady_clock::now();
        // set_viewdirs[_vr] over exactly (gx, gy) items: the kernels test
        // `> sample_x`, so a larger range would write past the records.
        int32_t bx = 256;
        while (gx % bx) bx /= 2;
        const F3 vp{view[0], view[1], view[2], 0}, tp{up[0], up[1], up[2], 0},
            rp{right[0], right[1], right[2], 0};
        if (height < 0) {
            int32_t shift = (int32_t)right[0] * ssx;  // trace_runner.cpp:132-140
            int32_t kx = sample_x * ssy, ky = sample_y * ssy;
            void *a[] = {(void *)&vp, &shift, &kx, &ky, &d_initial};
            if (!ok(hipModuleLaunchKernel(fn[1], gx / bx, gy, 1, bx, 1, 1, 0, nullptr, a, nullptr),
                    "set_viewdirs_vr"))
                break;
        } else {
            void *a[] = {(void *)&vp, (void *)&tp, (void *)&rp, &width, &height,
                         (void *)&gx, (void *)&gy, &d_initial};
            if (!ok(hipModuleLaunchKernel(fn[0], gx / bx, gy, 1, bx, 1, 1, 0, nullptr, a, nullptr),
                    "set_viewdirs"))
                break;
        }
        if (!ok(hipDeviceSynchronize(), "set_viewdirs")) break;

        int32_t sizes[10];
        auto put = [&]() {
            return ok(hipMemcpy(d_sizes, sizes, 40, hipMemcpyHostToDevice), "write sizes");
        };
        auto get = [&]() {
            return ok(hipMemcpy(sizes, d_sizes, 40, hipMemcpyDeviceToHost), "read sizes");
        };
        // ND_RANGE of trace_runner.cpp:183-186: (n/256 + 1) groups of 256.
        auto nd = [&](hipFunction_t f, int32_t n, void **a, const char *what) {
            return ok(hipModuleLaunchKernel(f, (unsigned)(n / 256 + 1), 1, 1, 256, 1, 1, 0,
                                            nullptr, a, nullptr),
                      what);
        };
        const F3 bg{background[0], background[1], background[2], 0};
        int32_t n_nodes32 = (int32_t)n_nodes, normals_size = (int32_t)n_normals,
                texcoords_size = (int32_t)n_texcoords, lights_size = (int32_t)n_lights;

        // The first intersect (sample_n = sample) and the cached S1 list.
        std::memset(sizes, 0, sizeof(sizes));
        sizes[0] = R;
        if (!put()) break;
        {
            int32_t sample_n = sample;
            void *a[] = {&d_sizes, &d_initial, &d_cache, &d_stage[1], &d_points, &d_mesh,
                         &d_leaf, &d_nodes, &n_nodes32, &d_results, (void *)&bg, &sample_n};
            if (!nd(fn[2], R, a, "kdtree_intersect")) break;
        }
        if (!get()) break;
        if (!ok(hipMemcpy(d_cache, d_stage[1], 4 * (size_t)R, hipMemcpyDeviceToDevice), "cache s1"))
            break;
        const int32_t initial_s1 = sizes[1];
        if (primary_geometry) {
            std::vector<Unit> recs(R);
            std::vector<int32_t> hits(initial_s1 > 0 ? initial_s1 : 1);
            if (!ok(hipMemcpy(recs.data(), d_initial, sizeof(Unit) * (size_t)R,
                              hipMemcpyDeviceToHost),
                    "read records") ||
                !ok(hipMemcpy(hits.data(), d_cache, 4 * (size_t)initial_s1,
                              hipMemcpyDeviceToHost),
                    "read s1"))
                break;
            for (size_t k = 0; k < 4 * (size_t)R; ++k) primary_geometry[k] = -1;
            for (int32_t h = 0; h < initial_s1; ++h) {
                const int32_t ray = hits[h];
                if (ray < 0 || ray >= R) continue;
                std::memcpy(primary_geometry + 4 * (size_t)ray, recs[ray].geometry, 16);
            }
        }

        bool fine = true;
        for (int32_t ii = 0; ii < sample && fine; ++ii) {
            std::memset(sizes, 0, sizeof(sizes));
            sizes[1] = initial_s1;
            fine = put() &&
                   ok(hipMemcpy(d_data, d_initial, sizeof(Unit) * (size_t)R,
                                hipMemcpyDeviceToDevice),
                      "restore records") &&
                   ok(hipMemcpy(d_stage[1], d_cache, 4 * (size_t)R, hipMemcpyDeviceToDevice),
                      "restore s1");
            for (int32_t i = 0; i < depth && fine; ++i) {
                if (s0_sizes) s0_sizes[(size_t)ii * depth + i] = sizes[0];
                if (sizes[0]) {
                    int32_t one = 1;
                    void *a[] = {&d_sizes, &d_data, &d_stage[0], &d_stage[1], &d_points, &d_mesh,
                                 &d_leaf, &d_nodes, &n_nodes32, &d_results, (void *)&bg, &one};
                    fine = nd(fn[2], sizes[0], a, "kdtree_intersect");
                }
                fine = fine && ok(hipDeviceSynchronize(), "kdtree_intersect") && get();
                if (!fine) break;
                sizes[0] = 0;
                if (sizes[1]) {
                    void *a[] = {&d_sizes,     &d_data,       &d_stage[1],     &d_results,
                                 &d_stage[2],  &d_stage[3],   &d_stage[4],     &d_stage[5],
                                 &d_points,    &d_normals,    &normals_size,   &d_texcoords,
                                 &texcoords_size, &d_mats,    &no_texture,     &disable_diffuse,
                                 &d_seeds};
                    fine = nd(fn[3], sizes[1], a, "s1_run");
                }
                fine = fine && ok(hipDeviceSynchronize(), "s1_run") && get();
                if (!fine) break;
                sizes[0] = sizes[1] = 0;
                if (!(fine = put())) break;
                if (i == depth - 1) break;
                void *ar[] = {&d_sizes, &d_data, &d_stage[2], &d_stage[0]};
                void *as[] = {&d_sizes, &d_data, &d_stage[3], &d_stage[0], &d_seeds};
                void *ad[] = {&d_sizes, &d_data, &d_stage[4], &d_stage[0], &d_seeds};
                void *al[] = {&d_sizes,  &d_data,      &d_stage[5], &d_stage[0],
                              &d_lights, &lights_size, &d_points,   &d_seeds};
                void **args[4] = {ar, as, ad, al};
                for (int k = 0; k < 4 && fine; ++k)
                    if (sizes[k + 2]) fine = nd(fn[4 + k], sizes[k + 2], args[k], names[4 + k]);
                fine = fine && ok(hipDeviceSynchronize(), "s2") && get();
                if (!fine) break;
                sizes[2] = sizes[3] = sizes[4] = sizes[5] = 0;
                fine = put();
            }
        }
        if (!fine) break;
        if (loop_seconds)
            *loop_seconds =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t_loop).count();
        if (!ok(hipMemcpy(results, d_results, 12 * (size_t)R, hipMemcpyDeviceToHost),
                "download results"))
            break;
        rc = 0;
    } while (false);
    for (void *p : keep) (void)hipFree(p);
    if (mod) (void)hipModuleUnload(mod);
    return rc;
}

int ref_kdtree_intersect(const char *co_path, const float *points, int64_t n_points,
                         const int32_t *mesh, int64_t n_tris, const void *nodes, int64_t n_nodes,
                         const int32_t *leaf, int64_t n_leaf, const float *org, const float *dir,
                         int32_t n, int32_t *geometry, float *t) {
    return ref_kdtree_intersect_timed(co_path, points, n_points, mesh, n_tris, nodes, n_nodes,
                                      leaf, n_leaf, org, dir, n, geometry, t, nullptr);
}

}  // extern "C"
