// Launcher of the reference's own kdtree_intersect (compiled from kernel.cl
// by oracle/build_ref.sh) -- TEST INFRASTRUCTURE.
//
// Reproduces the first intersect of TraceRunner::run (trace_runner.cpp:
// 192-206): every ray is a 144-byte unit_data record (types.h.cl:8-17) with
// orig_id = index, strength = 1, start_p and in_dir set; v_s0 = 0..n-1;
// sample_n = 1 and a black background; 256-wide groups over
// (n/256 + 1)*256 items.  Returns the geometry int4 and intersect_number the
// kernel writes into each record (left at -1 / -42 on a miss).
#include <hip/hip_runtime.h>

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

int ref_kdtree_intersect(const char *co_path, const float *points, int64_t n_points,
                         const int32_t *mesh, int64_t n_tris, const void *nodes, int64_t n_nodes,
                         const int32_t *leaf, int64_t n_leaf, const float *org, const float *dir,
                         int32_t n, int32_t *geometry, float *t) {
    return ref_kdtree_intersect_timed(co_path, points, n_points, mesh, n_tris, nodes, n_nodes,
                                      leaf, n_leaf, org, dir, n, geometry, t, nullptr);
}

}  // extern "C"
