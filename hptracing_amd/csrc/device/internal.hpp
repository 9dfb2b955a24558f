// Internal state shared by the C ABI translation units of libhptracing.so:
// error plumbing, device allocations, kernel timing and the scene object.
#pragma once

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

namespace hpt {

inline thread_local std::string g_error;

struct Fail {
    int code;
    std::string msg;
};

[[noreturn]] inline void fail(int code, const std::string &msg) { throw Fail{code, msg}; }

inline void hip_check(hipError_t e, const char *what) {
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
inline void upload(DevMem &m, const T *src, size_t count, int64_t &total) {
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
    DevMem hit_id, hit_t, accum, rgba, views, segs, live, prim_count, totals, pass, passes;
    // The work counter of every cast this workspace's owner launches.  All
    // of them run on the owner's one stream, one after another, and each
    // launch's last wave re-arms the counter, so one counter serves them
    // all and no other caller can reach it.
    DevMem counter;
    int64_t path_capacity = 0;
    // rays: camera rays (and accumulator entries per slice); paths: live
    // paths of one launch (up to kPassSlices per camera ray).
    void ensure(int64_t rays, int64_t pixels, int n_passes, int depth, int n_views,
                int64_t paths) {
        if (rays > capacity) {
            const size_t n = (size_t)rays;
            prim_org.reserve(16 * n);
            prim_dir.reserve(16 * n);
            prim_thr.reserve(16 * n);
            prim_id.reserve(4 * n);
            prim_t.reserve(4 * n);
            prim_hit.reserve(8 * n);
            prim_list.reserve(4 * n);
            accum.reserve(12 * n * hpt::kPassSlices);
            capacity = rays;
        }
        if (paths > path_capacity) {
            const size_t n = (size_t)paths;
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
            path_capacity = paths;
        }
        if (!counter.p) {
            const WorkCounter zero{0, 0};
            counter.reserve(sizeof(WorkCounter));
            hip_check(hipMemcpy(counter.p, &zero, sizeof(zero), hipMemcpyHostToDevice),
                      "work counter");
        }
        rgba.reserve(4 * (size_t)std::max<int64_t>(pixels, 1));
        views.reserve(sizeof(hpt::ViewParams) * (size_t)std::max(n_views, 1));
        segs.reserve(sizeof(hpt::Segment) * (size_t)std::max(n_passes * n_views, 1));
        live.reserve(sizeof(int) * (size_t)n_passes * (depth + 1));
        prim_count.reserve(sizeof(int) * (size_t)std::max(n_views, 1));
        totals.reserve(4 * sizeof(unsigned long long));
        pass.reserve(sizeof(int));
        passes.reserve(sizeof(hpt::PassDesc) * (size_t)std::max(n_passes, 1));
    }
};

// The work counter of one hpt_cast call, from the stream-ordered allocator
// and released on the same stream: casts running concurrently on other
// streams (or threads) never share a counter.  Cleared by a fill kernel, so
// a captured graph holds no memset node.
struct StreamCounter {
    hipStream_t st;
    WorkCounter *p = nullptr;
    explicit StreamCounter(hipStream_t s) : st(s) {
        void *q = nullptr;
        hip_check(hipMallocAsync(&q, sizeof(WorkCounter), st), "hipMallocAsync");
        p = static_cast<WorkCounter *>(q);
        launch_zero_counter(p, st);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) {
            (void)hipFreeAsync(p, st);
            p = nullptr;
            hip_check(e, "k_zero_counter");
        }
    }
    StreamCounter(const StreamCounter &) = delete;
    StreamCounter &operator=(const StreamCounter &) = delete;
    ~StreamCounter() {
        if (p) (void)hipFreeAsync(p, st);
    }
};

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

// The launch settings of a scene, copied out under its settings lock by
// callers that launch without holding the scene's render lock.
struct LaunchForm {
    DevScene dev;
    int cast_grid, shade_grid, sort_rays;
};

}  // namespace hpt

// Locking: `mu` is held by renders, batch renders and the render autotuner
// for their whole run (they share `ws` and `stream`).  The launch settings
// (dev, cast_grid, shade_grid, sort_rays) change only under `mu` and
// `tune_mu` together, so a holder of `mu` reads them directly, and every
// other caller (hpt_cast, hpt_cast_noalloc, the server's worker, the cast
// autotuner) takes a copy with form(), which holds only `tune_mu` and so
// never waits for a render.
struct hpt_scene {
    int device = 0;
    hpt::DevScene dev{};
    hpt::DevMem nodes, tris, tri_recs, leaf_ids, points, normals, texcoords, mesh, materials,
        lights, textures;
    int64_t device_bytes = 0;
    int64_t n_nodes = 0, n_entries = 0, n_triangles = 0, n_textures = 0;
    int cast_grid = 0, shade_grid = 0, n_cus = 1;
    int sort_rays = 0;                 // hpt_tuning.sort_rays
    float box_lo[3] = {0, 0, 0}, box_hi[3] = {0, 0, 0};  // root node's box
    std::mutex mu;
    mutable std::mutex tune_mu;
    hpt::Workspace ws;  // hpt_render / hpt_batch_render / hpt_autotune, under mu
    hipStream_t stream = nullptr;
    hpt::LaunchForm form() const {
        std::lock_guard<std::mutex> lock(tune_mu);
        return hpt::LaunchForm{dev, cast_grid, shade_grid, sort_rays};
    }
};

namespace hpt {
// Smallest batch that sort_rays orders: below it the sort costs more than
// the locality gains.
constexpr int64_t kSortMinRays = 65536;
// hpt_cast of a batch in Morton order of its origins (sort.hip); results
// in the caller's order.  Buffers from the stream-ordered allocator.
void cast_sorted(const hpt_scene *s, const LaunchForm &f, const float *org3, const float *dir3,
                 int64_t n, int *out_id, float *out_t, hipStream_t st);
// The same with the caller's buffers, each of n entries (keys and idx
// twice: the radix sort's double buffer), temp_bytes >= sort_temp_bytes(n)
// and a work counter cleared to zero.
struct SortBuffers {
    uint32_t *keys[2];
    int32_t *idx[2];
    float4 *org4, *dir4;
    int *ids;
    float *t;
    void *temp;
    size_t temp_bytes;
    WorkCounter *wc;
};
size_t sort_temp_bytes(int64_t n, hipStream_t st);
void cast_sorted_into(const hpt_scene *s, const LaunchForm &f, const float *org3,
                      const float *dir3, int64_t n, int *out_id, float *out_t,
                      const SortBuffers &b, hipStream_t st);
}  // namespace hpt
