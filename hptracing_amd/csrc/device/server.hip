// Continuous-batching render server (hpt_server_*).  No reference
// counterpart: the reference renders one view per blocking
// TraceRunner::run call; its GUI keeps one scene resident across renders
// (src/hptracing.cpp:124-130, 249), which is the starting point here.
//
// One worker thread owns a stream and a device arena sized by the caller's
// memory budget.  It loops over "macro passes":
//   1. admit queued requests in arrival order while their camera rays fit in
//      a free range of the arena (first fit); a request larger than the whole
//      arena is refused at submit time;
//   2. for newly admitted requests: camera rays, primary cast, primary-hit
//      cache (their first intersect with sample_n = spp);
//   3. the next kServerPasses sample passes of every active request that
//      still has samples to take: the same expand / cast / shade launches
//      as hpt_render, over the union of their live paths;
//   4. requests whose last sample finished are resolved, copied to the
//      caller's buffers and released; waiters are woken.  A progressive
//      request (HPT_SERVE_PROGRESSIVE) is also resolved after every macro
//      pass that advanced it, into an image the caller reads with
//      hpt_server_peek: the progressive preview the reference's GUI gives by
//      re-rendering its view (src/hptracing.cpp:60-83, 124-130).
// Requests arriving while others run join at the next macro pass.  Paths
// carry their own random streams, so a request's image does not depend on
// what it shared the GPU with.
#include <condition_variable>
#include <deque>
#include <map>
#include <thread>

#include "internal.hpp"

using namespace hpt;

namespace {

constexpr int kMaxSlots = 1024;
// Sample passes per macro pass.  Fewer than a render's launch group
// (kPassSlices), so that a request arriving while others render waits less
// for the next admission; pass p still adds into slice p % kPassSlices, so
// results are the same bits as a render's.
constexpr int kServerPasses = 8;
static_assert(kServerPasses <= kPassSlices, "passes in flight must use distinct slices");
// Device bytes per camera-ray slot of the arena: primary org/dir/thr (48),
// primary hit id/t (8), cached hit (8), hit list (4), RGBA (4), the folded
// accumulator (12), an accumulator per pass slice (12 each), and per pass
// of a macro pass two path buffers (2 x 56) and hit id/t of the live paths
// (8).
constexpr int64_t kBytesPerRay =
    48 + 8 + 8 + 4 + 4 + 12 + kPassSlices * 12 + kServerPasses * (112 + 8);

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct Request {
    int64_t id = 0;
    hpt_view v{};
    uint8_t *rgba = nullptr;
    float *accum = nullptr;
    int64_t sx = 0, sy = 0, R = 0, pixels = 0;
    int64_t base = -1;  // first arena slot
    int slot = -1;      // index into the device view table
    int n1 = 0;
    int next_pass = 0;
    int status = 0;     // 0 queued, 1 active, 2 done, 3 failed
    bool progressive = false;
    std::vector<uint8_t> stage;    // worker side: the image just resolved
    std::vector<uint8_t> preview;  // under the server mutex: the latest image
    int published = 0;             // passes in `preview` (under the mutex)
    int peekers = 0;               // hpt_server_peek calls inside (under the mutex)
    double t_submit = 0, t_admit = 0, t_done = 0;
    std::string error;
};

// First-fit allocator of ranges in [0, capacity).
struct RangeAlloc {
    int64_t capacity = 0;
    std::map<int64_t, int64_t> free_;  // start -> length
    void init(int64_t cap) {
        capacity = cap;
        free_.clear();
        if (cap > 0) free_[0] = cap;
    }
    int64_t take(int64_t n) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                const int64_t start = it->first, len = it->second;
                free_.erase(it);
                if (len > n) free_[start + n] = len - n;
                return start;
            }
        }
        return -1;
    }
    void give(int64_t start, int64_t n) {
        auto it = free_.emplace(start, n).first;
        auto next = std::next(it);
        if (next != free_.end() && start + n == next->first) {
            it->second += next->second;
            free_.erase(next);
        }
        if (it != free_.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == it->first) {
                prev->second += it->second;
                free_.erase(it);
            }
        }
    }
};

}  // namespace

struct hpt_server {
    hpt_scene *scene = nullptr;
    int64_t capacity = 0;   // camera-ray slots
    int max_depth = HPT_MAX_DEPTH;
    hipStream_t stream = nullptr;
    Workspace ws;           // arena: indexed by slot of the camera-ray range
    DevMem folded;          // the slices of a request summed, per camera ray
    RangeAlloc alloc;
    std::vector<int> free_slots;

    std::mutex mu;
    std::condition_variable wake, done_cv;
    std::deque<Request *> queue;
    std::vector<Request *> active;
    std::map<int64_t, std::unique_ptr<Request>> requests;
    int64_t next_id = 1;
    bool stopping = false;
    std::string fatal;
    std::thread worker;

    // statistics
    int64_t completed = 0, macro_passes = 0, rays_cast = 0;
    double busy_s = 0;

    void run();
    void admit_locked(std::vector<Request *> &admitted);
    void start_requests(const LaunchForm &f, const std::vector<Request *> &admitted);
    void one_pass(const LaunchForm &f, std::vector<Request *> &finished,
                  std::vector<Request *> &progressed);
    void resolve(Request *r, int samples);
    void publish(const std::vector<Request *> &progressed);
    void finish(const std::vector<Request *> &finished);
};

void hpt_server::admit_locked(std::vector<Request *> &admitted) {
    while (!queue.empty() && !free_slots.empty()) {
        Request *r = queue.front();
        const int64_t base = alloc.take(r->R);
        if (base < 0) break;  // arrival order: the head waits for space
        queue.pop_front();
        r->base = base;
        r->slot = free_slots.back();
        free_slots.pop_back();
        r->status = 1;
        r->t_admit = now_s();
        active.push_back(r);
        admitted.push_back(r);
    }
}

void hpt_server::start_requests(const LaunchForm &f, const std::vector<Request *> &admitted) {
    hipStream_t st = stream;
    ViewParams *d_views = ws.views.as<ViewParams>();
    int *prim_count = ws.prim_count.as<int>();
    Paths prim{ws.prim_org.as<float4>(), ws.prim_dir.as<float4>(), ws.prim_thr.as<float4>(), nullptr};
    for (Request *r : admitted) {
        ViewParams vp{};
        const hpt_view &v = r->v;
        for (int k = 0; k < 3; ++k) {
            vp.view[k] = v.view[k];
            vp.background[k] = v.background[k];
            vp.up[k] = v.up[k];
            vp.right[k] = v.right[k];
        }
        vp.disable_diffuse = v.disable_diffuse;
        vp.depth = v.depth;
        vp.sample = v.sample;
        vp.sx = (int32_t)r->sx;
        vp.sy = (int32_t)r->sy;
        vp.vr = v.height < 0 ? 1 : 0;
        vp.vr_shift = (int32_t)v.right[0] * v.supersample_x;
        vp.vr_sample_x = v.sample_x * v.supersample_y;
        vp.vr_sample_y = v.sample_y * v.supersample_y;
        vp.width = v.width;
        vp.height = v.height;
        vp.seed = v.seed;
        vp.ray_base = r->base;
        vp.accum_stride = capacity;
        hip_check(hipMemcpyAsync(d_views + r->slot, &vp, sizeof(vp), hipMemcpyHostToDevice, st),
                  "upload view");
        hip_check(hipMemsetAsync(prim_count + r->slot, 0, sizeof(int), st), "memset");
        launch_camera(d_views, r->slot, r->R, prim, st);
        launch_cast(f.dev, nullptr, nullptr, prim.org + r->base, prim.dir + r->base, CountRef{},
                    r->R, ws.counter.as<WorkCounter>(), ws.prim_id.as<int>() + r->base,
                    ws.prim_t.as<float>() + r->base, nullptr, nullptr, f.cast_grid, st);
        launch_primary(d_views, r->slot, r->R, ws.prim_id.as<int>(), ws.prim_t.as<float>(),
                       ws.accum.as<float>(), ws.prim_hit.as<int2>(), ws.prim_list.as<int>(),
                       prim_count, st);
    }
    std::vector<int> counts(kMaxSlots);
    hip_check(hipMemcpyAsync(counts.data(), prim_count, sizeof(int) * kMaxSlots,
                             hipMemcpyDeviceToHost, st), "read counts");
    hip_check(hipStreamSynchronize(st), "primary pass");
    for (Request *r : admitted) {
        r->n1 = counts[r->slot];
        rays_cast += r->R;
    }
}

void hpt_server::one_pass(const LaunchForm &f, std::vector<Request *> &finished,
                          std::vector<Request *> &progressed) {
    hipStream_t st = stream;
    std::vector<Segment> segs;
    PassDesc pd{0, 0, 0};
    int depth = 1;
    for (Request *r : active) {
        if (r->next_pass >= r->v.sample) continue;
        const int end = std::min(r->next_pass + kServerPasses, r->v.sample);
        if (r->n1 > 0) {
            for (int p = r->next_pass; p < end; ++p) {
                segs.push_back(Segment{r->slot, p, r->base, r->n1, pd.total});
                pd.total += r->n1;
                ++pd.n_segs;
            }
            depth = std::max(depth, r->v.depth);
        }
        r->next_pass = end;
        if (r->next_pass >= r->v.sample) finished.push_back(r);
        else if (r->progressive) progressed.push_back(r);
    }
    if (pd.total > 0) {
        hip_check(hipMemcpyAsync(ws.segs.p, segs.data(), sizeof(Segment) * segs.size(),
                                 hipMemcpyHostToDevice, st), "upload segments");
        hip_check(hipMemcpyAsync(ws.passes.p, &pd, sizeof(pd), hipMemcpyHostToDevice, st),
                  "upload pass");
        hip_check(hipMemsetAsync(ws.pass.p, 0, sizeof(int), st), "memset");
        hip_check(hipMemsetAsync(ws.live.p, 0, sizeof(int) * (HPT_MAX_DEPTH + 1), st), "memset");
        int *pass = ws.pass.as<int>();
        int *live = ws.live.as<int>();
        const int stride = depth + 1;
        ViewParams *d_views = ws.views.as<ViewParams>();
        Paths cur{ws.a_org.as<float4>(), ws.a_dir.as<float4>(), ws.a_thr.as<float4>(), ws.a_seed.as<uint64_t>()};
        Paths nxt{ws.b_org.as<float4>(), ws.b_dir.as<float4>(), ws.b_thr.as<float4>(), ws.b_seed.as<uint64_t>()};
        launch_expand(d_views, ws.segs.as<Segment>(), ws.passes.as<PassDesc>(), pass, pd.total,
                      ws.prim_list.as<int>(), ws.prim_hit.as<int2>(), ws.prim_dir.as<float4>(), cur,
                      ws.hit_id.as<int>(), ws.hit_t.as<float>(), CountRef{live, pass, stride, 0}, st);
        for (int i = 0; i < depth; ++i) {
            if (i > 0)
                launch_cast(f.dev, nullptr, nullptr, cur.org, cur.dir,
                            CountRef{live, pass, stride, i}, 0, ws.counter.as<WorkCounter>(),
                            ws.hit_id.as<int>(), ws.hit_t.as<float>(), nullptr, nullptr,
                            f.cast_grid, st);
            launch_shade(f.dev, d_views, i, CountRef{live, pass, stride, i}, cur,
                         ws.hit_id.as<int>(), ws.hit_t.as<float>(), nxt,
                         CountRef{live, pass, stride, i + 1}, ws.accum.as<float>(),
                         f.shade_grid, st);
            std::swap(cur, nxt);
        }
        std::vector<int> lv(depth + 1);
        hip_check(hipMemcpyAsync(lv.data(), live, sizeof(int) * lv.size(), hipMemcpyDeviceToHost, st),
                  "read live");
        hip_check(hipStreamSynchronize(st), "pass");
        for (int i = 1; i < depth; ++i) rays_cast += lv[i];
    }
    ++macro_passes;
}

// Resolve request r as an image of `samples` passes into the arena: the
// background of the missed camera rays for that many samples, the slices
// summed into the folded accumulator (the slices stay as they are, so the
// request can go on), then the tone map.  After `samples` passes the image
// equals hpt_render's for the same view with sample = samples.
void hpt_server::resolve(Request *r, int samples) {
    hipStream_t st = stream;
    ViewParams *d_views = ws.views.as<ViewParams>();
    const hpt_view &v = r->v;
    if (r->progressive)
        launch_misses(d_views, r->slot, r->R, ws.prim_id.as<int>(), samples, ws.accum.as<float>(),
                      st);
    launch_fold(d_views, r->slot, r->R, ws.accum.as<float>(), folded.as<float>(), st);
    uchar4 *img = ws.rgba.as<uchar4>() + r->base;  // pixels <= camera rays
    launch_resolve(d_views, r->slot, v.sample_x, v.sample_y, v.supersample_x, v.supersample_y,
                   samples, v.brightness, folded.as<float>(), img, st);
}

void hpt_server::publish(const std::vector<Request *> &progressed) {
    if (progressed.empty()) return;
    for (Request *r : progressed) {
        resolve(r, r->next_pass);
        hip_check(hipMemcpyAsync(r->stage.data(), ws.rgba.as<uchar4>() + r->base,
                                 4 * (size_t)r->pixels, hipMemcpyDeviceToHost, stream),
                  "read preview");
    }
    hip_check(hipStreamSynchronize(stream), "preview");
    std::lock_guard<std::mutex> lock(mu);
    for (Request *r : progressed) {
        r->preview = r->stage;
        r->published = r->next_pass;
    }
    done_cv.notify_all();
}

void hpt_server::finish(const std::vector<Request *> &finished) {
    if (finished.empty()) return;
    hipStream_t st = stream;
    for (Request *r : finished) {
        resolve(r, r->v.sample);
        const uchar4 *img = ws.rgba.as<uchar4>() + r->base;
        if (r->rgba)
            hip_check(hipMemcpyAsync(r->rgba, img, 4 * (size_t)r->pixels, hipMemcpyDeviceToHost, st),
                      "read image");
        if (r->progressive)
            hip_check(hipMemcpyAsync(r->stage.data(), img, 4 * (size_t)r->pixels,
                                     hipMemcpyDeviceToHost, st), "read preview");
        if (r->accum)
            hip_check(hipMemcpyAsync(r->accum, folded.as<float>() + 3 * r->base, 12 * (size_t)r->R,
                                     hipMemcpyDeviceToHost, st), "read accum");
    }
    hip_check(hipStreamSynchronize(st), "resolve");
}

void hpt_server::run() {
    hip_check(hipSetDevice(scene->device), "hipSetDevice");
    for (;;) {
        std::vector<Request *> admitted, finished, progressed;
        {
            std::unique_lock<std::mutex> lock(mu);
            wake.wait(lock, [&] { return stopping || !queue.empty() || !active.empty(); });
            if (stopping && queue.empty() && active.empty()) return;
            admit_locked(admitted);
            if (active.empty()) {
                // Head of the queue cannot fit although nothing runs: only
                // possible if it exceeds the arena, which submit refuses.
                continue;
            }
        }
        const double t0 = now_s();
        try {
            // The scene's launch settings as they are now, for this macro
            // pass (settings never change a result).
            const LaunchForm f = scene->form();
            if (!admitted.empty()) start_requests(f, admitted);
            one_pass(f, finished, progressed);
            publish(progressed);
            finish(finished);
        } catch (const Fail &e) {
            std::lock_guard<std::mutex> lock(mu);
            fatal = e.msg;
            for (Request *r : active) {
                r->status = 3;
                r->error = e.msg;
                r->t_done = now_s();
            }
            active.clear();
            done_cv.notify_all();
            return;
        }
        busy_s += now_s() - t0;
        std::lock_guard<std::mutex> lock(mu);
        const double t = now_s();
        for (Request *r : finished) {
            r->status = 2;
            r->t_done = t;
            if (r->progressive) {
                r->preview = r->stage;
                r->published = r->v.sample;
            }
            alloc.give(r->base, r->R);
            free_slots.push_back(r->slot);
            active.erase(std::find(active.begin(), active.end(), r));
            ++completed;
        }
        if (!finished.empty()) done_cv.notify_all();
    }
}

extern "C" {

int hpt_server_create(hpt_scene *scene, const hpt_server_opts *opts, hpt_server **out) {
    return guarded([&] {
        if (!scene || !out) fail(HPT_ERR_INVALID, "null argument");
        *out = nullptr;
        int64_t budget = opts ? opts->device_budget_bytes : 0;
        if (budget <= 0) budget = (int64_t)4 << 30;
        std::unique_ptr<hpt_server> sv(new hpt_server());
        sv->scene = scene;
        sv->capacity = budget / kBytesPerRay;
        if (sv->capacity < 64) fail(HPT_ERR_BUDGET, "device budget too small for any request");
        if (sv->capacity > INT32_MAX / kPassSlices) sv->capacity = INT32_MAX / kPassSlices;
        sv->ws.ensure(sv->capacity, sv->capacity, 1, HPT_MAX_DEPTH, kMaxSlots,
                      sv->capacity * kServerPasses);
        sv->ws.segs.reserve(sizeof(Segment) * kMaxSlots * kServerPasses);
        sv->folded.reserve(12 * (size_t)sv->capacity);
        sv->alloc.init(sv->capacity);
        for (int k = kMaxSlots - 1; k >= 0; --k) sv->free_slots.push_back(k);
        hip_check(hipStreamCreateWithFlags(&sv->stream, hipStreamNonBlocking), "hipStreamCreate");
        hpt_server *raw = sv.get();
        sv->worker = std::thread([raw] {
            try {
                raw->run();
            } catch (const Fail &e) {
                std::lock_guard<std::mutex> lock(raw->mu);
                raw->fatal = e.msg;
                raw->done_cv.notify_all();
            }
        });
        *out = sv.release();
    });
}

int hpt_server_submit(hpt_server *sv, const hpt_view *v, uint8_t *rgba, float *accum,
                      int64_t *id) {
    return hpt_server_submit_ex(sv, v, rgba, accum, 0u, id);
}

int hpt_server_submit_ex(hpt_server *sv, const hpt_view *v, uint8_t *rgba, float *accum,
                         unsigned flags, int64_t *id) {
    return guarded([&] {
        if (!sv || !v || !id) fail(HPT_ERR_INVALID, "null argument");
        if (flags & ~HPT_SERVE_PROGRESSIVE) fail(HPT_ERR_INVALID, "unknown flags");
        if (v->sample_x <= 0 || v->sample_y <= 0 || v->supersample_x <= 0 || v->supersample_y <= 0)
            fail(HPT_ERR_INVALID, "image and supersample sizes must be positive");
        if (v->sample <= 0 || v->depth <= 0 || v->depth > HPT_MAX_DEPTH)
            fail(HPT_ERR_INVALID, "sample must be >= 1 and depth in [1, 64]");
        if (v->height < 0 && v->supersample_x != v->supersample_y)
            fail(HPT_ERR_INVALID, "the equirectangular camera needs supersample_x == supersample_y");
        std::unique_ptr<Request> r(new Request());
        r->v = *v;
        r->rgba = rgba;
        r->accum = accum;
        r->sx = (int64_t)v->sample_x * v->supersample_x;
        r->sy = (int64_t)v->sample_y * v->supersample_y;
        r->R = r->sx * r->sy;
        r->pixels = (int64_t)v->sample_x * v->sample_y;
        if (flags & HPT_SERVE_PROGRESSIVE) {
            r->progressive = true;
            r->stage.assign(4 * (size_t)r->pixels, 0);
        }
        if (r->R > sv->capacity)
            fail(HPT_ERR_BUDGET, "request needs " + std::to_string(r->R * kBytesPerRay) +
                                     " device bytes, more than the server's budget");
        std::lock_guard<std::mutex> lock(sv->mu);
        if (!sv->fatal.empty()) fail(HPT_ERR_STATE, "server stopped: " + sv->fatal);
        if (sv->stopping) fail(HPT_ERR_STATE, "server is shutting down");
        r->id = sv->next_id++;
        r->t_submit = now_s();
        *id = r->id;
        sv->queue.push_back(r.get());
        sv->requests[r->id] = std::move(r);
        sv->wake.notify_one();
    });
}

static void fill_info(const Request &r, hpt_request_info *info) {
    if (!info) return;
    std::memset(info, 0, sizeof(*info));
    info->id = r.id;
    info->status = r.status;
    info->queue_ms = r.t_admit > 0 ? (r.t_admit - r.t_submit) * 1e3 : 0.0;
    info->latency_ms = r.t_done > 0 ? (r.t_done - r.t_submit) * 1e3 : 0.0;
    info->camera_rays = r.R;
}

int hpt_server_wait(hpt_server *sv, int64_t id, double timeout_s, hpt_request_info *info) {
    int code = HPT_OK;
    int rc = guarded([&] {
        if (!sv) fail(HPT_ERR_INVALID, "null server");
        std::unique_lock<std::mutex> lock(sv->mu);
        auto it = sv->requests.find(id);
        if (it == sv->requests.end()) fail(HPT_ERR_INVALID, "unknown request id");
        Request *r = it->second.get();
        auto ready = [&] { return r->status >= 2 || !sv->fatal.empty(); };
        if (timeout_s < 0) {
            sv->done_cv.wait(lock, ready);
        } else if (!sv->done_cv.wait_for(lock, std::chrono::duration<double>(timeout_s), ready)) {
            fill_info(*r, info);
            code = HPT_ERR_TIMEOUT;
            return;
        }
        fill_info(*r, info);
        if (r->status == 3 || (r->status < 2 && !sv->fatal.empty()))
            fail(HPT_ERR_HIP, "request failed: " + (r->error.empty() ? sv->fatal : r->error));
        sv->done_cv.wait(lock, [&] { return r->peekers == 0; });
        sv->requests.erase(it);  // results delivered; forget the request
    });
    if (rc == HPT_OK && code != HPT_OK) {
        g_error = "request not finished before the timeout";
        return code;
    }
    return rc;
}

int hpt_server_peek(hpt_server *sv, int64_t id, int32_t min_passes, double timeout_s,
                    uint8_t *rgba, int32_t *passes) {
    int code = HPT_OK;
    int rc = guarded([&] {
        if (!sv) fail(HPT_ERR_INVALID, "null server");
        std::unique_lock<std::mutex> lock(sv->mu);
        auto it = sv->requests.find(id);
        if (it == sv->requests.end()) fail(HPT_ERR_INVALID, "unknown request id");
        Request *r = it->second.get();
        if (!r->progressive) fail(HPT_ERR_INVALID, "not a progressive request");
        auto ready = [&] {
            return r->published >= min_passes || r->status >= 2 || !sv->fatal.empty();
        };
        // hpt_server_wait does not retire the request while a peek is inside.
        ++r->peekers;
        if (timeout_s < 0) {
            sv->done_cv.wait(lock, ready);
        } else if (!sv->done_cv.wait_for(lock, std::chrono::duration<double>(timeout_s), ready)) {
            code = HPT_ERR_TIMEOUT;
        }
        if (--r->peekers == 0) sv->done_cv.notify_all();
        if (r->status == 3 || (r->status < 2 && !sv->fatal.empty()))
            fail(HPT_ERR_HIP, "request failed: " + (r->error.empty() ? sv->fatal : r->error));
        if (passes) *passes = r->published;
        if (rgba && r->published > 0) std::memcpy(rgba, r->preview.data(), r->preview.size());
    });
    if (rc == HPT_OK && code != HPT_OK) {
        g_error = "fewer passes published before the timeout";
        return code;
    }
    return rc;
}

int hpt_server_stats_get(hpt_server *sv, hpt_server_stats *out) {
    return guarded([&] {
        if (!sv || !out) fail(HPT_ERR_INVALID, "null argument");
        std::lock_guard<std::mutex> lock(sv->mu);
        std::memset(out, 0, sizeof(*out));
        out->queued = (int64_t)sv->queue.size();
        out->active = (int64_t)sv->active.size();
        out->completed = sv->completed;
        out->macro_passes = sv->macro_passes;
        out->rays_cast = sv->rays_cast;
        out->busy_seconds = sv->busy_s;
        out->capacity_rays = sv->capacity;
    });
}

int hpt_server_destroy(hpt_server *sv) {
    return guarded([&] {
        if (!sv) return;
        {
            std::lock_guard<std::mutex> lock(sv->mu);
            sv->stopping = true;
            sv->wake.notify_all();
        }
        if (sv->worker.joinable()) sv->worker.join();
        if (sv->stream) {
            (void)hipStreamSynchronize(sv->stream);
            (void)hipStreamDestroy(sv->stream);
        }
        delete sv;
    });
}

}  // extern "C"
