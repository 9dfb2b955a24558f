// C ABI of the GPU side (hpt_scene_*, hpt_cast*, hpt_render), declared in
// include/hptracing.h.  Replaces the OpenCL host layer hp::TraceRunner
// (src/hp/trace_runner.cpp): scene upload in hpt_scene_create, the render
// loop of TraceRunner::run in hpt_render.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "hptracing.h"
#include "kernels.hpp"
#include "launch.hpp"

namespace {

thread_local std::string g_error;

struct Fail {
    int code;
    std::string msg;
};

[[noreturn]] void fail(int code, const std::string &msg) { throw Fail{code, msg}; }

void hip_check(hipError_t e, const char *what) {
    if (e != hipSuccess) fail(HPT_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

template <class F>
int guarded(F &&f) {
    try {
        g_error.clear();
        f();
        return HPT_OK;
    } catch (const Fail &e) {
        g_error = e.msg;
        return e.code;
    } catch (const std::bad_alloc &) {
        g_error = "host allocation failed";
        return HPT_ERR_NOMEM;
    } catch (const std::exception &e) {
        g_error = e.what();
        return HPT_ERR_INVALID;
    }
}

// Device allocation owned by a scene or a workspace.
struct DevMem {
    void *p = nullptr;
    size_t bytes = 0;
    DevMem() = default;
    DevMem(const DevMem &) = delete;
    DevMem &operator=(const DevMem &) = delete;
    ~DevMem() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void reserve(size_t n) {
        if (n <= bytes) return;
        release();
        if (n == 0) return;
        hipError_t e = hipMalloc(&p, n);
        if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
            fail(HPT_ERR_NOMEM, "device allocation of " + std::to_string(n) + " bytes failed");
        hip_check(e, "hipMalloc");
        bytes = n;
    }
    template <class T>
    T *as() const {
        return static_cast<T *>(p);
    }
};

template <class T>
void upload(DevMem &m, const T *src, size_t count, int64_t &total) {
    size_t n = sizeof(T) * std::max<size_t>(count, 1);
    m.reserve(n);
    if (count) hip_check(hipMemcpy(m.p, src, sizeof(T) * count, hipMemcpyHostToDevice), "upload");
    total += (int64_t)n;
}

// Render buffers for up to `capacity` camera rays, grown on demand.
struct Workspace {
    int64_t capacity = 0;
    DevMem prim_org, prim_dir, prim_thr, prim_id, prim_t, prim_hit, prim_list;
    DevMem a_org, a_dir, a_thr, a_seed, b_org, b_dir, b_thr, b_seed;
    DevMem hit_id, hit_t, accum, rgba, views, segs, live, prim_count, totals;
    void ensure(int64_t rays, int64_t pixels, int passes, int depth, int n_views) {
        if (rays > capacity) {
            const size_t n = (size_t)rays;
            prim_org.reserve(16 * n);
            prim_dir.reserve(16 * n);
            prim_thr.reserve(16 * n);
            prim_id.reserve(4 * n);
            prim_t.reserve(4 * n);
            prim_hit.reserve(8 * n);
            prim_list.reserve(4 * n);
            a_org.reserve(16 * n);
            a_dir.reserve(16 * n);
            a_thr.reserve(16 * n);
            a_seed.reserve(8 * n);
            b_org.reserve(16 * n);
            b_dir.reserve(16 * n);
            b_thr.reserve(16 * n);
            b_seed.reserve(8 * n);
            hit_id.reserve(4 * n);
            hit_t.reserve(4 * n);
            accum.reserve(12 * n);
            capacity = rays;
        }
        rgba.reserve(4 * (size_t)std::max<int64_t>(pixels, 1));
        views.reserve(sizeof(hpt::ViewParams) * (size_t)std::max(n_views, 1));
        segs.reserve(sizeof(hpt::Segment) * (size_t)std::max(passes * n_views, 1));
        live.reserve(sizeof(int) * (size_t)passes * (depth + 1));
        prim_count.reserve(sizeof(int) * (size_t)std::max(n_views, 1));
        totals.reserve(4 * sizeof(unsigned long long));
    }
};

constexpr int kCounterRing = 256;

// Device time of a series of launches on one stream, from HIP event pairs
// recorded around each launch (read after the stream is synchronized).
struct KernelTimer {
    bool on;
    hipStream_t st;
    std::vector<hipEvent_t> ev;
    KernelTimer(bool enabled, hipStream_t s) : on(enabled), st(s) {}
    ~KernelTimer() {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
    void mark() {
        hipEvent_t e;
        hip_check(hipEventCreate(&e), "hipEventCreate");
        ev.push_back(e);
        hip_check(hipEventRecord(e, st), "hipEventRecord");
    }
    void begin() {
        if (on) mark();
    }
    void end() {
        if (on) mark();
    }
    int64_t count() const { return (int64_t)ev.size() / 2; }
    double total_ms() const {
        double t = 0;
        for (size_t i = 0; i + 1 < ev.size(); i += 2) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, ev[i], ev[i + 1]) == hipSuccess) t += ms;
        }
        return t;
    }
};

}  // namespace

struct hpt_scene {
    int device = 0;
    hpt::DevScene dev{};
    DevMem nodes, tris, points, normals, texcoords, mesh, materials, lights, textures, counters;
    int64_t device_bytes = 0;
    int64_t n_nodes = 0, n_entries = 0, n_triangles = 0, n_textures = 0;
    int cast_grid = 0, shade_grid = 0;
    std::atomic<int> next_counter{0};
    std::mutex mu;
    Workspace ws;
    hipStream_t stream = nullptr;

    hpt::WorkCounter *counter() {
        const int k = next_counter.fetch_add(1) % kCounterRing;
        return counters.as<hpt::WorkCounter>() + k;
    }
};

namespace {

// Escape index of every node: where the reference's walk continues when it
// does not descend (sibling, else the parent's escape; past the root = n).
std::vector<int32_t> escape_indices(const hpt_kd_node *nodes, int64_t n) {
    std::vector<int32_t> esc(n, -2);
    for (int64_t i = 0; i < n; ++i) {
        // Follow parents until a node with a sibling or the root.
        std::vector<int64_t> chain;
        int64_t j = i;
        int32_t target = -2;
        while (true) {
            if (esc[j] != -2) {
                target = esc[j];
                break;
            }
            chain.push_back(j);
            if (chain.size() > (size_t)n + 1) fail(HPT_ERR_SCENE, "KD node links form a cycle");
            const hpt_kd_node &nd = nodes[j];
            if (nd.sibling >= 0) {
                target = nd.sibling;
                break;
            }
            if (nd.parent < 0) {
                target = (int32_t)n;
                break;
            }
            j = nd.parent;
        }
        for (int64_t c : chain) esc[c] = target;
    }
    return esc;
}

}  // namespace

extern "C" {

const char *hpt_last_error(void) { return g_error.c_str(); }

int hpt_scene_create(const hpt_scene_desc *d, hpt_scene **out) {
    return guarded([&] {
        if (!d || !out) fail(HPT_ERR_INVALID, "null argument");
        *out = nullptr;
        if (d->n_points <= 0 || d->n_triangles <= 0 || d->n_materials <= 0 || d->n_nodes <= 0)
            fail(HPT_ERR_SCENE, "scene needs points, triangles, materials and KD nodes");
        if (d->n_nodes > INT32_MAX / 2 || d->n_points > INT32_MAX || d->n_triangles > INT32_MAX)
            fail(HPT_ERR_INVALID, "scene too large for 32-bit indices");
        if (d->n_lights != 0 && d->n_lights != HPT_LIGHT_TABLE_SIZE)
            fail(HPT_ERR_SCENE, "light table must have 0 or 256 entries");
        const int64_t np = d->n_points, nt = d->n_triangles, nm = d->n_materials;
        // Validate every index the kernels will follow.
        int32_t max_vertex = -1;
        bool textured = false;
        for (int64_t m = 0; m < nm; ++m) {
            const hpt_material &mt = d->materials[m];
            if (mt.diffuse_texture_id >= 0 || mt.ambient_texture_id >= 0) textured = true;
            if (mt.diffuse_texture_id >= d->n_textures || mt.ambient_texture_id >= d->n_textures)
                fail(HPT_ERR_SCENE, "material references a missing texture layer");
        }
        for (int64_t i = 0; i < nt; ++i) {
            const int32_t *t = d->mesh + 4 * i;
            for (int k = 0; k < 3; ++k) {
                if (t[k] < 0 || t[k] >= np) fail(HPT_ERR_SCENE, "triangle references a missing vertex");
                max_vertex = std::max(max_vertex, t[k]);
            }
            if (t[3] < 0 || t[3] >= nm) fail(HPT_ERR_SCENE, "triangle references a missing material");
        }
        if (d->n_normals != 0 && max_vertex >= d->n_normals)
            fail(HPT_ERR_SCENE, "normal stream shorter than the vertex indices it is read with");
        if (textured && d->n_texcoords != 0 && max_vertex >= d->n_texcoords)
            fail(HPT_ERR_SCENE, "texcoord stream shorter than the vertex indices it is read with");
        for (int64_t i = 0; i < d->n_lights; ++i)
            for (int k = 0; k < 3; ++k)
                if (d->lights[4 * i + k] < 0 || d->lights[4 * i + k] >= np)
                    fail(HPT_ERR_SCENE, "light entry references a missing vertex");

        // Repack the KD tree: 32-byte nodes with escape indices and 48-byte
        // triangle records in leaf order.
        const int64_t nn = d->n_nodes;
        const hpt_kd_node *nodes = d->nodes;
        std::vector<int32_t> esc = escape_indices(nodes, nn);
        std::vector<float> pnodes(8 * nn);
        std::vector<float> recs;
        int64_t entries = 0;
        for (int64_t i = 0; i < nn; ++i) {
            const hpt_kd_node &nd = nodes[i];
            if (nd.child >= 0 && nd.child != i + 1)
                fail(HPT_ERR_SCENE, "KD nodes are not in pre-order (child != index + 1)");
            // The walk must move forward at every step so that it ends.
            if ((nd.child < 0 && esc[i] != i + 1) || esc[i] <= i || esc[i] > nn)
                fail(HPT_ERR_SCENE, "KD sibling/parent links do not follow the pre-order");
            if (nd.child < 0 && nd.data >= 0) {
                if (nd.data >= d->n_leaf_data) fail(HPT_ERR_SCENE, "leaf offset out of range");
                const int32_t cnt = d->leaf_data[nd.data];
                if (cnt < 0 || (int64_t)nd.data + 1 + cnt > d->n_leaf_data)
                    fail(HPT_ERR_SCENE, "leaf list out of range");
                entries += cnt;
            }
        }
        if (entries > INT32_MAX / 3) fail(HPT_ERR_INVALID, "too many leaf entries");
        recs.resize(12 * std::max<int64_t>(entries, 1));
        int64_t rec = 0;
        for (int64_t i = 0; i < nn; ++i) {
            const hpt_kd_node &nd = nodes[i];
            int32_t a, b;
            if (nd.child >= 0) {
                a = esc[i];
                b = -1;
            } else if (nd.data >= 0) {
                const int32_t cnt = d->leaf_data[nd.data];
                a = cnt;
                b = (int32_t)rec;
                for (int32_t k = 0; k < cnt; ++k) {
                    const int32_t tri = d->leaf_data[nd.data + 1 + k];
                    if (tri < 0 || tri >= nt) fail(HPT_ERR_SCENE, "leaf references a missing triangle");
                    const int32_t *t = d->mesh + 4 * (int64_t)tri;
                    const float *pa = d->points + 4 * (int64_t)t[0];
                    const float *pb = d->points + 4 * (int64_t)t[1];
                    const float *pc = d->points + 4 * (int64_t)t[2];
                    float *r = &recs[12 * rec];
                    r[0] = pa[0]; r[1] = pa[1]; r[2] = pa[2];
                    std::memcpy(&r[3], &tri, 4);
                    r[4] = pb[0] - pa[0]; r[5] = pb[1] - pa[1]; r[6] = pb[2] - pa[2]; r[7] = 0.f;
                    r[8] = pc[0] - pa[0]; r[9] = pc[1] - pa[1]; r[10] = pc[2] - pa[2]; r[11] = 0.f;
                    ++rec;
                }
            } else {
                a = 0;  // empty leaf
                b = 0;
            }
            float *p = &pnodes[8 * i];
            p[0] = nd.box_start[0]; p[1] = nd.box_start[1]; p[2] = nd.box_start[2];
            std::memcpy(&p[3], &a, 4);
            p[4] = nd.box_end[0]; p[5] = nd.box_end[1]; p[6] = nd.box_end[2];
            std::memcpy(&p[7], &b, 4);
        }

        std::unique_ptr<hpt_scene> s(new hpt_scene());
        hip_check(hipGetDevice(&s->device), "hipGetDevice");
        int64_t total = 0;
        upload(s->nodes, pnodes.data(), pnodes.size(), total);
        upload(s->tris, recs.data(), recs.size(), total);
        upload(s->points, d->points, 4 * (size_t)np, total);
        upload(s->normals, d->normals, 4 * (size_t)d->n_normals, total);
        upload(s->texcoords, d->texcoords, 2 * (size_t)d->n_texcoords, total);
        upload(s->mesh, d->mesh, 4 * (size_t)nt, total);
        upload(s->materials, d->materials, (size_t)nm, total);
        upload(s->lights, d->lights, 4 * (size_t)d->n_lights, total);
        upload(s->textures, d->textures, (size_t)d->n_textures * HPT_TEXTURE_SIZE * HPT_TEXTURE_SIZE * 4, total);
        std::vector<hpt::WorkCounter> zero(kCounterRing, hpt::WorkCounter{0, 0});
        upload(s->counters, zero.data(), zero.size(), total);
        s->device_bytes = total;
        s->n_nodes = nn;
        s->n_entries = entries;
        s->n_triangles = nt;
        s->n_textures = d->n_textures;

        hpt::DevScene &ds = s->dev;
        ds.nodes = s->nodes.as<float4>();
        ds.n_nodes = (int32_t)nn;
        ds.tris = s->tris.as<float4>();
        ds.points = s->points.as<float4>();
        ds.normals = s->normals.as<float4>();
        ds.n_normals = (int32_t)d->n_normals;
        ds.texcoords = s->texcoords.as<float2>();
        ds.n_texcoords = (int32_t)d->n_texcoords;
        ds.mesh = s->mesh.as<int4>();
        ds.materials = s->materials.as<hpt_material>();
        ds.lights = s->lights.as<int4>();
        ds.n_lights = (int32_t)d->n_lights;
        ds.textures = s->textures.as<uchar4>();
        ds.n_textures = (int32_t)d->n_textures;

        int cus = 0;
        hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device),
                  "hipDeviceGetAttribute");
        s->cast_grid = std::max(1, cus * std::max(1, hpt::occupancy_blocks_cast()));
        s->shade_grid = std::max(1, cus * std::max(1, hpt::occupancy_blocks_shade()));
        hip_check(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking), "hipStreamCreate");
        *out = s.release();
    });
}

int hpt_scene_destroy(hpt_scene *s) {
    return guarded([&] {
        if (!s) return;
        if (s->stream) {
            (void)hipStreamSynchronize(s->stream);
            (void)hipStreamDestroy(s->stream);
        }
        delete s;
    });
}

int hpt_scene_stats_get(const hpt_scene *s, hpt_scene_stats *out) {
    return guarded([&] {
        if (!s || !out) fail(HPT_ERR_INVALID, "null argument");
        out->n_nodes = s->n_nodes;
        out->n_leaf_entries = s->n_entries;
        out->n_triangles = s->n_triangles;
        out->n_textures = s->n_textures;
        out->device_bytes = s->device_bytes;
    });
}

static int cast_impl(hpt_scene *s, const float *org, const float *dir, int64_t n, int32_t *tri_id,
                     float *t, int32_t *counts, void *stream) {
    return guarded([&] {
        if (!s) fail(HPT_ERR_INVALID, "null scene");
        if (n < 0 || n > INT32_MAX - 64) fail(HPT_ERR_INVALID, "ray count out of range");
        if (n == 0) return;
        if (!org || !dir || !tri_id || !t) fail(HPT_ERR_INVALID, "null buffer");
        hipStream_t st = stream ? (hipStream_t)stream : s->stream;
        hpt::launch_cast(s->dev, org, dir, nullptr, nullptr, nullptr, n, s->counter(), tri_id, t,
                         counts, nullptr, s->cast_grid, st);
        hip_check(hipGetLastError(), "k_cast launch");
    });
}

int hpt_cast(hpt_scene *s, const float *org, const float *dir, int64_t n, int32_t *tri_id,
             float *t, void *stream) {
    return cast_impl(s, org, dir, n, tri_id, t, nullptr, stream);
}

int hpt_cast_counted(hpt_scene *s, const float *org, const float *dir, int64_t n,
                     int32_t *tri_id, float *t, int32_t *counts, void *stream) {
    if (!counts) {
        g_error = "null counts buffer";
        return HPT_ERR_INVALID;
    }
    return cast_impl(s, org, dir, n, tri_id, t, counts, stream);
}

int hpt_render(hpt_scene *s, const hpt_view *v, uint8_t *rgba, float *accum_out,
               hpt_render_stats *stats) {
    return hpt_render_ex(s, v, rgba, accum_out, stats, 0u);
}

int hpt_render_ex(hpt_scene *s, const hpt_view *v, uint8_t *rgba, float *accum_out,
                  hpt_render_stats *stats, unsigned flags) {
    return guarded([&] {
        if (!s || !v) fail(HPT_ERR_INVALID, "null argument");
        if (v->sample_x <= 0 || v->sample_y <= 0 || v->supersample_x <= 0 || v->supersample_y <= 0)
            fail(HPT_ERR_INVALID, "image and supersample sizes must be positive");
        if (v->sample <= 0 || v->depth <= 0 || v->depth > HPT_MAX_DEPTH)
            fail(HPT_ERR_INVALID, "sample must be >= 1 and depth in [1, 64]");
        const bool vr = v->height < 0;
        if (vr && v->supersample_x != v->supersample_y)
            fail(HPT_ERR_INVALID, "the equirectangular camera needs supersample_x == supersample_y");
        const int64_t sx = (int64_t)v->sample_x * v->supersample_x;
        const int64_t sy = (int64_t)v->sample_y * v->supersample_y;
        const int64_t R = sx * sy;
        if (R > INT32_MAX / 2) fail(HPT_ERR_INVALID, "too many camera rays");
        const int64_t pixels = (int64_t)v->sample_x * v->sample_y;
        const auto t0 = std::chrono::steady_clock::now();

        std::lock_guard<std::mutex> lock(s->mu);
        hipStream_t st = s->stream;
        Workspace &ws = s->ws;
        ws.ensure(R, pixels, v->sample, v->depth, 1);

        hpt::ViewParams vp{};
        for (int k = 0; k < 3; ++k) {
            vp.view[k] = v->view[k];
            vp.background[k] = v->background[k];
            vp.up[k] = v->up[k];
            vp.right[k] = v->right[k];
        }
        vp.disable_diffuse = v->disable_diffuse;
        vp.depth = v->depth;
        vp.sample = v->sample;
        vp.sx = (int32_t)sx;
        vp.sy = (int32_t)sy;
        vp.vr = vr ? 1 : 0;
        // trace_runner.cpp:131-140: shift = int(right.x) * ssx; the kernel's
        // sample_x / sample_y arguments are W*ssy and H*ssy.
        vp.vr_shift = (int32_t)v->right[0] * v->supersample_x;
        vp.vr_sample_x = v->sample_x * v->supersample_y;
        vp.vr_sample_y = v->sample_y * v->supersample_y;
        vp.width = v->width;
        vp.height = v->height;
        vp.seed = v->seed;
        vp.ray_base = 0;
        hpt::ViewParams *d_views = ws.views.as<hpt::ViewParams>();
        hip_check(hipMemcpyAsync(d_views, &vp, sizeof(vp), hipMemcpyHostToDevice, st), "upload view");

        hpt::Paths prim{ws.prim_org.as<float4>(), ws.prim_dir.as<float4>(), ws.prim_thr.as<float4>(), nullptr};
        hpt::Paths pa{ws.a_org.as<float4>(), ws.a_dir.as<float4>(), ws.a_thr.as<float4>(), ws.a_seed.as<uint64_t>()};
        hpt::Paths pb{ws.b_org.as<float4>(), ws.b_dir.as<float4>(), ws.b_thr.as<float4>(), ws.b_seed.as<uint64_t>()};
        int *prim_count = ws.prim_count.as<int>();
        int *live = ws.live.as<int>();
        float *accum = ws.accum.as<float>();
        hip_check(hipMemsetAsync(prim_count, 0, sizeof(int), st), "memset");
        hip_check(hipMemsetAsync(live, 0, sizeof(int) * (size_t)v->sample * (v->depth + 1), st), "memset");

        // Camera rays, the primary cast (once, as the reference caches it)
        // and the primary-hit list.
        hpt::launch_camera(d_views, 0, R, prim, st);
        const bool timed = (flags & HPT_RENDER_TIME_KERNELS) != 0;
        unsigned long long *totals =
            (flags & HPT_RENDER_COUNT_WORK) ? ws.totals.as<unsigned long long>() : nullptr;
        if (totals) hip_check(hipMemsetAsync(totals, 0, 3 * sizeof(unsigned long long), st), "memset");
        KernelTimer cast_timer(timed, st), shade_timer(timed, st);
        cast_timer.begin();
        hpt::launch_cast(s->dev, nullptr, nullptr, prim.org, prim.dir, nullptr, R, s->counter(),
                         ws.prim_id.as<int>(), ws.prim_t.as<float>(), nullptr, totals,
                         s->cast_grid, st);
        cast_timer.end();
        hpt::launch_primary(d_views, 0, R, ws.prim_id.as<int>(), ws.prim_t.as<float>(), accum,
                            ws.prim_hit.as<int2>(), ws.prim_list.as<int>(), prim_count, st);
        int n1 = 0;
        hip_check(hipMemcpyAsync(&n1, prim_count, sizeof(int), hipMemcpyDeviceToHost, st), "read count");
        hip_check(hipStreamSynchronize(st), "sync");

        std::vector<hpt::Segment> segs(v->sample);
        for (int p = 0; p < v->sample; ++p) segs[p] = hpt::Segment{0, p, 0, n1, 0};
        hpt::Segment *d_segs = ws.segs.as<hpt::Segment>();
        hip_check(hipMemcpyAsync(d_segs, segs.data(), sizeof(hpt::Segment) * segs.size(),
                                 hipMemcpyHostToDevice, st), "upload segments");
        for (int p = 0; p < v->sample && n1 > 0; ++p) {
            int *lv = live + (size_t)p * (v->depth + 1);
            hpt::Paths cur = pa, nxt = pb;
            hpt::launch_expand(d_views, d_segs + p, 1, n1, ws.prim_list.as<int>(),
                               ws.prim_hit.as<int2>(), prim.dir, cur, ws.hit_id.as<int>(),
                               ws.hit_t.as<float>(), lv, st);
            for (int i = 0; i < v->depth; ++i) {
                if (i > 0) {
                    cast_timer.begin();
                    hpt::launch_cast(s->dev, nullptr, nullptr, cur.org, cur.dir, lv + i, 0,
                                     s->counter(), ws.hit_id.as<int>(), ws.hit_t.as<float>(),
                                     nullptr, totals, s->cast_grid, st);
                    cast_timer.end();
                }
                shade_timer.begin();
                hpt::launch_shade(s->dev, d_views, i, lv + i, cur, ws.hit_id.as<int>(),
                                  ws.hit_t.as<float>(), nxt, lv + i + 1, accum, s->shade_grid, st);
                shade_timer.end();
                std::swap(cur, nxt);
            }
        }
        hip_check(hipGetLastError(), "render launches");
        if (rgba) {
            hpt::launch_resolve(d_views, 0, v->sample_x, v->sample_y, v->supersample_x,
                                v->supersample_y, v->brightness, accum, ws.rgba.as<uchar4>(), st);
            hip_check(hipMemcpyAsync(rgba, ws.rgba.p, 4 * (size_t)pixels, hipMemcpyDeviceToHost, st),
                      "read image");
        }
        if (accum_out)
            hip_check(hipMemcpyAsync(accum_out, accum, 12 * (size_t)R, hipMemcpyDeviceToHost, st),
                      "read accum");
        std::vector<int> lv_host((size_t)v->sample * (v->depth + 1));
        hip_check(hipMemcpyAsync(lv_host.data(), live, sizeof(int) * lv_host.size(),
                                 hipMemcpyDeviceToHost, st), "read live counts");
        unsigned long long work[3] = {0, 0, 0};
        if (totals)
            hip_check(hipMemcpyAsync(work, totals, sizeof(work), hipMemcpyDeviceToHost, st),
                      "read work counters");
        hip_check(hipStreamSynchronize(st), "render");
        if (stats) {
            std::memset(stats, 0, sizeof(*stats));
            stats->cast_ms = cast_timer.total_ms();
            stats->cast_launches = cast_timer.count();
            stats->shade_ms = shade_timer.total_ms();
            stats->shade_launches = shade_timer.count();
            stats->node_visits = (int64_t)work[0];
            stats->leaf_visits = (int64_t)work[1];
            stats->tri_tests = (int64_t)work[2];
            stats->primary_rays = R;
            stats->live_per_depth[0] = R;
            int64_t cast = R;
            for (int p = 0; p < v->sample; ++p)
                for (int i = 1; i < v->depth; ++i) {
                    const int64_t c = lv_host[(size_t)p * (v->depth + 1) + i];
                    stats->live_per_depth[i] += c;
                    cast += c;
                }
            stats->rays_cast = cast;
            stats->seconds =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
    });
}

}  // extern "C"
