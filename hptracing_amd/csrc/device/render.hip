// C ABI of the render entry points (hpt_render*, hpt_batch_render): the
// loop of TraceRunner::run (src/hp/trace_runner.cpp:116-356) on the GPU,
// generalised to several views that share every launch.
//
// Per render call:
//   1. camera rays of every view (k_camera), one primary cast over all of
//      them (k_cast), then per view the primary-hit cache and the
//      background of the misses (k_primary) -- the reference's first
//      intersect with sample_n = spp (trace_runner.cpp:192-213);
//   2. one host read of the primary-hit counts, from which the host writes
//      a table of passes: pass p covers every view with spp > p;
//   3. per pass: k_expand restores the cached hits of those views into the
//      path buffer, then `depth` levels of k_cast + k_shade with the live
//      counts kept on the device (no host round trip), then k_next_pass.
//      The pass sequence is captured once in a hipGraph and replayed;
//   4. k_resolve per view.
#include "internal.hpp"

using namespace hpt;

namespace {

struct Job {
    const hpt_view *v;
    int64_t sx, sy, R, pixels;
    int64_t ray_base, pix_base;
    int n1 = 0;
};

void validate_view(const hpt_view *v) {
    if (!v) fail(HPT_ERR_INVALID, "null view");
    if (v->sample_x <= 0 || v->sample_y <= 0 || v->supersample_x <= 0 || v->supersample_y <= 0)
        fail(HPT_ERR_INVALID, "image and supersample sizes must be positive");
    if (v->sample <= 0 || v->depth <= 0 || v->depth > HPT_MAX_DEPTH)
        fail(HPT_ERR_INVALID, "sample must be >= 1 and depth in [1, 64]");
    if (v->height < 0 && v->supersample_x != v->supersample_y)
        fail(HPT_ERR_INVALID, "the equirectangular camera needs supersample_x == supersample_y");
    const int64_t R = (int64_t)v->sample_x * v->supersample_x * v->sample_y * v->supersample_y;
    if (R > INT32_MAX / kPassSlices) fail(HPT_ERR_INVALID, "too many camera rays");
}

ViewParams params_of(const hpt_view &v, int64_t sx, int64_t sy, int64_t ray_base,
                     int64_t accum_stride) {
    ViewParams vp{};
    for (int k = 0; k < 3; ++k) {
        vp.view[k] = v.view[k];
        vp.background[k] = v.background[k];
        vp.up[k] = v.up[k];
        vp.right[k] = v.right[k];
    }
    vp.disable_diffuse = v.disable_diffuse;
    vp.depth = v.depth;
    vp.sample = v.sample;
    vp.sx = (int32_t)sx;
    vp.sy = (int32_t)sy;
    vp.vr = v.height < 0 ? 1 : 0;
    // trace_runner.cpp:131-140: shift = int(right.x) * ssx; the kernel's
    // sample_x / sample_y arguments are W*ssy and H*ssy.
    vp.vr_shift = (int32_t)v.right[0] * v.supersample_x;
    vp.vr_sample_x = v.sample_x * v.supersample_y;
    vp.vr_sample_y = v.sample_y * v.supersample_y;
    vp.width = v.width;
    vp.height = v.height;
    vp.seed = v.seed;
    vp.ray_base = ray_base;
    vp.accum_stride = accum_stride;
    return vp;
}

// Launches of one pass: expand, `depth` levels of cast + shade, next pass.
struct PassLauncher {
    hpt_scene *s;
    Workspace &ws;
    ViewParams *d_views;
    int depth;
    int64_t max_total;
    unsigned long long *totals;
    KernelTimer *cast_timer, *shade_timer;

    void enqueue(hipStream_t st) {
        int *pass = ws.pass.as<int>();
        int *live = ws.live.as<int>();
        const int stride = depth + 1;
        Paths cur{ws.a_org.as<float4>(), ws.a_dir.as<float4>(), ws.a_thr.as<float4>(), ws.a_seed.as<uint64_t>()};
        Paths nxt{ws.b_org.as<float4>(), ws.b_dir.as<float4>(), ws.b_thr.as<float4>(), ws.b_seed.as<uint64_t>()};
        launch_expand(d_views, ws.segs.as<Segment>(), ws.passes.as<PassDesc>(), pass, max_total,
                      ws.prim_list.as<int>(), ws.prim_hit.as<int2>(), ws.prim_dir.as<float4>(),
                      cur, ws.hit_id.as<int>(), ws.hit_t.as<float>(),
                      CountRef{live, pass, stride, 0}, st);
        for (int i = 0; i < depth; ++i) {
            if (i > 0) {
                if (cast_timer) cast_timer->begin();
                launch_cast(s->dev, nullptr, nullptr, cur.org, cur.dir, CountRef{live, pass, stride, i},
                            0, ws.counter.as<WorkCounter>(), ws.hit_id.as<int>(),
                            ws.hit_t.as<float>(), nullptr, totals, s->cast_grid, st);
                if (cast_timer) cast_timer->end();
            }
            if (shade_timer) shade_timer->begin();
            launch_shade(s->dev, d_views, i, CountRef{live, pass, stride, i}, cur,
                         ws.hit_id.as<int>(), ws.hit_t.as<float>(), nxt,
                         CountRef{live, pass, stride, i + 1}, ws.accum.as<float>(), s->shade_grid,
                         st);
            if (shade_timer) shade_timer->end();
            std::swap(cur, nxt);
        }
        launch_next_pass(pass, st);
    }
};

void render_batch(hpt_scene *s, const hpt_view *views, int n, uint8_t *const *rgba,
                  float *const *accum_out, hpt_render_stats *stats, unsigned flags) {
    if (!s) fail(HPT_ERR_INVALID, "null scene");
    if (n <= 0 || !views) fail(HPT_ERR_INVALID, "no views");
    const auto t0 = std::chrono::steady_clock::now();
    std::vector<Job> jobs(n);
    int64_t rays = 0, pixels = 0;
    int max_spp = 0, max_depth = 0;
    for (int k = 0; k < n; ++k) {
        const hpt_view *v = &views[k];
        validate_view(v);
        Job &j = jobs[k];
        j.v = v;
        j.sx = (int64_t)v->sample_x * v->supersample_x;
        j.sy = (int64_t)v->sample_y * v->supersample_y;
        j.R = j.sx * j.sy;
        j.pixels = (int64_t)v->sample_x * v->sample_y;
        j.ray_base = rays;
        j.pix_base = pixels;
        rays += j.R;
        pixels += j.pixels;
        max_spp = std::max(max_spp, v->sample);
        max_depth = std::max(max_depth, v->depth);
    }
    // Accumulator entries (camera ray + slice * rays) are 32-bit.
    if (rays > INT32_MAX / kPassSlices) fail(HPT_ERR_INVALID, "too many camera rays in one batch");
    // Passes per launch sequence: kPassSlices, or one (the reference's
    // structure), and fewer when the path buffers of kPassSlices passes
    // (two ping-pong buffers of 56 B and a hit id/t of 8 B per path) would
    // pass 48 GiB -- an 8K image with 2x2 supersampling, say.  Pass p adds
    // into slice p % kPassSlices in any grouping, so the sums are the same.
    constexpr int64_t kPathBytes = 2 * 56 + 8;
    constexpr int64_t kPathBudget = (int64_t)48 << 30;
    int group = (flags & HPT_RENDER_PASS_AT_A_TIME) ? 1 : kPassSlices;
    group = (int)std::max<int64_t>(
        1, std::min<int64_t>(group, kPathBudget / std::max<int64_t>(1, rays * kPathBytes)));
    const int max_groups = (max_spp + group - 1) / group;

    std::lock_guard<std::mutex> lock(s->mu);
    hipStream_t st = s->stream;
    Workspace &ws = s->ws;
    ws.ensure(rays, pixels, max_groups, max_depth, n, 0);
    const bool timed = (flags & HPT_RENDER_TIME_KERNELS) != 0;
    const bool use_graph = !timed && !(flags & HPT_RENDER_NO_GRAPH);
    unsigned long long *totals =
        (flags & HPT_RENDER_COUNT_WORK) ? ws.totals.as<unsigned long long>() : nullptr;

    std::vector<ViewParams> vps(n);
    for (int k = 0; k < n; ++k)
        vps[k] = params_of(*jobs[k].v, jobs[k].sx, jobs[k].sy, jobs[k].ray_base, rays);
    ViewParams *d_views = ws.views.as<ViewParams>();
    hip_check(hipMemcpyAsync(d_views, vps.data(), sizeof(ViewParams) * n, hipMemcpyHostToDevice, st),
              "upload views");
    int *prim_count = ws.prim_count.as<int>();
    int *live = ws.live.as<int>();
    hip_check(hipMemsetAsync(prim_count, 0, sizeof(int) * n, st), "memset");
    hip_check(hipMemsetAsync(live, 0, sizeof(int) * (size_t)max_groups * (max_depth + 1), st), "memset");
    hip_check(hipMemsetAsync(ws.pass.p, 0, sizeof(int), st), "memset");
    if (totals) hip_check(hipMemsetAsync(totals, 0, 3 * sizeof(unsigned long long), st), "memset");

    KernelTimer cast_timer(timed, st), shade_timer(timed, st);
    Paths prim{ws.prim_org.as<float4>(), ws.prim_dir.as<float4>(), ws.prim_thr.as<float4>(), nullptr};
    for (int k = 0; k < n; ++k) launch_camera(d_views, k, jobs[k].R, prim, st);
    cast_timer.begin();
    launch_cast(s->dev, nullptr, nullptr, prim.org, prim.dir, CountRef{}, rays,
                ws.counter.as<WorkCounter>(), ws.prim_id.as<int>(), ws.prim_t.as<float>(), nullptr,
                totals, s->cast_grid, st);
    cast_timer.end();
    for (int k = 0; k < n; ++k)
        launch_primary(d_views, k, jobs[k].R, ws.prim_id.as<int>(), ws.prim_t.as<float>(),
                       ws.accum.as<float>(), ws.prim_hit.as<int2>(), ws.prim_list.as<int>(),
                       prim_count, st);
    std::vector<int> n1(n);
    hip_check(hipMemcpyAsync(n1.data(), prim_count, sizeof(int) * n, hipMemcpyDeviceToHost, st),
              "read counts");
    hip_check(hipStreamSynchronize(st), "primary pass");

    // Launch groups: group q runs passes q*group .. (q+1)*group-1 of every
    // view with that many samples, all in the same launches.
    std::vector<Segment> segs;
    std::vector<PassDesc> passes(max_groups);
    int64_t max_total = 0;
    for (int q = 0; q < max_groups; ++q) {
        PassDesc pd{(int32_t)segs.size(), 0, 0};
        for (int k = 0; k < n; ++k) {
            if (n1[k] == 0) continue;
            for (int p = q * group; p < (q + 1) * group && p < jobs[k].v->sample; ++p) {
                segs.push_back(Segment{k, p, jobs[k].ray_base, n1[k], pd.total});
                pd.total += n1[k];
                ++pd.n_segs;
            }
        }
        passes[q] = pd;
        max_total = std::max(max_total, pd.total);
    }
    int live_passes = 0;
    while (live_passes < max_groups && passes[live_passes].total > 0) ++live_passes;
    ws.ensure(rays, pixels, max_groups, max_depth, n, max_total);
    if (!segs.empty()) {
        ws.segs.reserve(sizeof(Segment) * segs.size());
        hip_check(hipMemcpyAsync(ws.segs.p, segs.data(), sizeof(Segment) * segs.size(),
                                 hipMemcpyHostToDevice, st), "upload segments");
    }
    hip_check(hipMemcpyAsync(ws.passes.p, passes.data(), sizeof(PassDesc) * passes.size(),
                             hipMemcpyHostToDevice, st), "upload passes");

    PassLauncher pl{s, ws, d_views, max_depth, max_total, totals,
                    timed ? &cast_timer : nullptr, timed ? &shade_timer : nullptr};
    if (live_passes > 0) {
        if (use_graph) {
            hipGraph_t graph = nullptr;
            hipGraphExec_t exec = nullptr;
            hip_check(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal), "begin capture");
            pl.enqueue(st);
            hip_check(hipStreamEndCapture(st, &graph), "end capture");
            hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            if (e != hipSuccess) {
                (void)hipGraphDestroy(graph);
                hip_check(e, "hipGraphInstantiate");
            }
            for (int p = 0; p < live_passes; ++p) {
                e = hipGraphLaunch(exec, st);
                if (e != hipSuccess) break;
            }
            (void)hipGraphExecDestroy(exec);
            (void)hipGraphDestroy(graph);
            hip_check(e, "hipGraphLaunch");
        } else {
            for (int p = 0; p < live_passes; ++p) pl.enqueue(st);
        }
    }
    hip_check(hipGetLastError(), "render launches");

    for (int k = 0; k < n; ++k)
        launch_fold(d_views, k, jobs[k].R, ws.accum.as<float>(), ws.accum.as<float>(), st);
    for (int k = 0; k < n; ++k) {
        const hpt_view *v = jobs[k].v;
        if (rgba && rgba[k]) {
            uchar4 *img = ws.rgba.as<uchar4>() + jobs[k].pix_base;
            launch_resolve(d_views, k, v->sample_x, v->sample_y, v->supersample_x, v->supersample_y,
                           v->sample, v->brightness, ws.accum.as<float>(), img, st);
            hip_check(hipMemcpyAsync(rgba[k], img, 4 * (size_t)jobs[k].pixels,
                                     hipMemcpyDeviceToHost, st), "read image");
        }
        if (accum_out && accum_out[k])
            hip_check(hipMemcpyAsync(accum_out[k], ws.accum.as<float>() + 3 * jobs[k].ray_base,
                                     12 * (size_t)jobs[k].R, hipMemcpyDeviceToHost, st),
                      "read accum");
    }
    std::vector<int> lv((size_t)max_groups * (max_depth + 1));
    hip_check(hipMemcpyAsync(lv.data(), live, sizeof(int) * lv.size(), hipMemcpyDeviceToHost, st),
              "read live counts");
    unsigned long long work[3] = {0, 0, 0};
    if (totals)
        hip_check(hipMemcpyAsync(work, totals, sizeof(work), hipMemcpyDeviceToHost, st),
                  "read work counters");
    hip_check(hipStreamSynchronize(st), "render");
    if (stats) {
        std::memset(stats, 0, sizeof(*stats));
        stats->primary_rays = rays;
        stats->live_per_depth[0] = rays;
        int64_t cast = rays;
        for (int p = 0; p < max_groups; ++p)
            for (int i = 1; i < max_depth; ++i) {
                const int64_t c = lv[(size_t)p * (max_depth + 1) + i];
                stats->live_per_depth[i] += c;
                cast += c;
            }
        stats->rays_cast = cast;
        stats->cast_ms = cast_timer.total_ms();
        stats->cast_launches = cast_timer.count();
        stats->shade_ms = shade_timer.total_ms();
        stats->shade_launches = shade_timer.count();
        stats->node_visits = (int64_t)work[0];
        stats->leaf_visits = (int64_t)work[1];
        stats->tri_tests = (int64_t)work[2];
        stats->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
}

// Coordinate descent over the scene's launch forms from its current
// setting: each knob in turn takes the value with the shortest cast time
// (measure(true), in ms) while the others are held, two sweeps; then, if
// `shade`, the shade grid by measure(false).  Leaves the best setting on
// the scene and returns it.  No setting changes a result.
template <class Measure>
hpt_tuning tune(hpt_scene *s, Measure measure, bool shade) {
    hpt_tuning base;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        base.fast_division = s->dev.fast_div;
        base.traversal = s->dev.traversal;
        base.refill_threshold = s->dev.refill;
        base.node_steps = s->dev.node_steps;
        base.fetch = s->dev.fetch;
        base.walk_min = s->dev.walk_min;
        base.sort_rays = s->sort_rays;
        base.cast_blocks_per_cu = s->cast_grid / s->n_cus;
        base.shade_blocks_per_cu = s->shade_grid / s->n_cus;
    }
    auto apply = [&](const hpt_tuning &t) {
        std::lock_guard<std::mutex> lock(s->mu);
        std::lock_guard<std::mutex> tlock(s->tune_mu);
        s->dev.fast_div = t.fast_division;
        s->dev.traversal = t.traversal;
        s->dev.refill = t.refill_threshold;
        s->dev.node_steps = t.node_steps;
        s->dev.fetch = t.fetch;
        s->dev.walk_min = t.walk_min;
        s->sort_rays = t.sort_rays;
        s->cast_grid = std::max(1, t.cast_blocks_per_cu) * s->n_cus;
        s->shade_grid = std::max(1, t.shade_blocks_per_cu) * s->n_cus;
    };
    auto timed = [&](const hpt_tuning &t, bool cast) {
        apply(t);
        return measure(cast);
    };
    timed(base, true);  // warm caches and code
    hpt_tuning best = base;
    double best_ms = timed(best, true);
    auto try_knob = [&](auto set, std::initializer_list<int> values) {
        const hpt_tuning start = best;
        for (int v : values) {
            hpt_tuning t = start;
            set(t, v);
            if (!valid_fetch(t.fetch) || !fetch_fits(t.fetch, s->dev)) continue;
            // A knob moves only for a gain above the run-to-run noise.
            const double ms = timed(t, true);
            if (ms < best_ms * 0.99) {
                best_ms = ms;
                best = t;
            }
        }
    };
    // The lane-refill knobs only matter for traversal 2.
    for (int sweep = 0; sweep < 2; ++sweep) {
        try_knob([](hpt_tuning &t, int v) { t.traversal = v; }, {0, 1, 2});
        if (best.traversal == 2) {
            try_knob([](hpt_tuning &t, int v) { t.fetch = v; },
                     {0, 1, 4, 5, 12, 13, 20, 21, 28, 29, 52, 60, 84, 92, 116, 124, 148,
                      149, 156, 157, 189, 220, 252, 413, 445, 669, 701, 1469,
                      1725, 3517, 3773});
            try_knob([](hpt_tuning &t, int v) { t.refill_threshold = v; },
                     {8, 16, 24, 32, 48, 64});
            try_knob([](hpt_tuning &t, int v) { t.node_steps = v; }, {8, 16, 24, 32, 48, 64});
            try_knob([](hpt_tuning &t, int v) { t.walk_min = v; }, {0, 16, 24, 32, 40});
        }
        try_knob([](hpt_tuning &t, int v) { t.fast_division = v; }, {0, 1});
        try_knob([](hpt_tuning &t, int v) { t.cast_blocks_per_cu = v; }, {3, 4, 5, 6, 7, 8});
        if (!shade) try_knob([](hpt_tuning &t, int v) { t.sort_rays = v; }, {0, 1});
    }
    if (shade) {
        const int socc = std::max(1, base.shade_blocks_per_cu);
        hpt_tuning t0 = best;
        t0.shade_blocks_per_cu = socc;
        double best_shade = timed(t0, false);
        int best_sbpc = socc;
        for (int sb : {std::max(1, socc / 2), std::min(64, socc * 2)}) {
            if (sb == socc) continue;
            hpt_tuning t = best;
            t.shade_blocks_per_cu = sb;
            const double ms = timed(t, false);
            if (ms < best_shade * 0.99) {
                best_shade = ms;
                best_sbpc = sb;
            }
        }
        best.shade_blocks_per_cu = best_sbpc;
    }
    apply(best);
    return best;
}

}  // namespace

extern "C" {

int hpt_render(hpt_scene *s, const hpt_view *v, uint8_t *rgba, float *accum_out,
               hpt_render_stats *stats) {
    return hpt_render_ex(s, v, rgba, accum_out, stats, 0u);
}

int hpt_render_ex(hpt_scene *s, const hpt_view *v, uint8_t *rgba, float *accum_out,
                  hpt_render_stats *stats, unsigned flags) {
    return guarded([&] {
        if (!s || !v) fail(HPT_ERR_INVALID, "null argument");
        render_batch(s, v, 1, &rgba, &accum_out, stats, flags);
    });
}

int hpt_batch_render(hpt_scene *s, const hpt_view *views, int n_views, uint8_t *const *rgba,
                     float *const *accum, hpt_render_stats *stats, unsigned flags) {
    return guarded([&] { render_batch(s, views, n_views, rgba, accum, stats, flags); });
}

int hpt_autotune(hpt_scene *s, const hpt_view *view, int max_rays, hpt_tuning *chosen) {
    return guarded([&] {
        if (!s || !view) fail(HPT_ERR_INVALID, "null argument");
        validate_view(view);
        // A reduced copy of the view: same camera and depth, the passes of
        // one launch group (a render casts up to kPassSlices passes per
        // launch, and the fill and tail of a launch depend on its size), the
        // image shrunk until camera rays x passes is at most max_rays.
        hpt_view v = *view;
        v.sample = std::min(view->sample, kPassSlices);
        if (max_rays <= 0) max_rays = 1 << 24;
        while ((int64_t)v.sample_x * v.supersample_x * v.sample_y * v.supersample_y * v.sample >
                   max_rays &&
               v.sample_x > 8 && v.sample_y > 8) {
            v.sample_x /= 2;
            v.sample_y /= 2;
        }
        auto measure = [&](bool cast) {
            hpt_render_stats st;
            double best = 1e30;
            for (int rep = 0; rep < 2; ++rep) {
                render_batch(s, &v, 1, nullptr, nullptr, &st, HPT_RENDER_TIME_KERNELS);
                best = std::min(best, cast ? st.cast_ms : st.shade_ms);
            }
            return best;
        };
        const hpt_tuning best = tune(s, measure, true);
        if (chosen) *chosen = best;
    });
}

int hpt_autotune_cast(hpt_scene *s, const float *org, const float *dir, int64_t n,
                      int64_t max_rays, hpt_tuning *chosen) {
    return guarded([&] {
        if (!s || !org || !dir) fail(HPT_ERR_INVALID, "null argument");
        if (n <= 0 || n > ((int64_t)1 << 30)) fail(HPT_ERR_INVALID, "ray count out of range");
        if (max_rays <= 0) max_rays = (int64_t)1 << 22;
        const int64_t m = std::min(n, max_rays);
        DevMem ids, ts;
        ids.reserve(4 * (size_t)m);
        ts.reserve(4 * (size_t)m);
        hipStream_t st = s->stream;
        auto measure = [&](bool) {
            double best = 1e30;
            for (int rep = 0; rep < 2; ++rep) {
                StreamCounter wc(st);
                KernelTimer timer(true, st);
                timer.begin();
                // as hpt_cast runs it, the sort of sort_rays included
                const LaunchForm f = s->form();
                if (f.sort_rays && m >= kSortMinRays)
                    cast_sorted(s, f, org, dir, m, ids.as<int>(), ts.as<float>(), st);
                else
                    launch_cast(f.dev, org, dir, nullptr, nullptr, CountRef{}, m, wc.p,
                                ids.as<int>(), ts.as<float>(), nullptr, nullptr, f.cast_grid,
                                st);
                timer.end();
                hip_check(hipStreamSynchronize(st), "autotune cast");
                best = std::min(best, timer.total_ms());
            }
            return best;
        };
        const hpt_tuning best = tune(s, measure, false);
        if (chosen) *chosen = best;
    });
}

}  // extern "C"
