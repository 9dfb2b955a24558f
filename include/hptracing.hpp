/*
 * hptracing.hpp - C++ host interface with the shape of the reference's
 * hp::KDTree / hp::TraceRunner (src/hp/scene/kdtree.h:54-59,
 * src/hp/trace_runner.h:21-66), implemented over the C ABI in hptracing.h.
 *
 * A caller of the reference switches by including this header and using
 * namespace hpt instead of hp: the constructors and run() keep their
 * parameter lists, meaning and defaults; failures throw hpt::Error (the
 * reference throws cl::Error out of run()).  The scene stays resident on
 * the device for the TraceRunner's lifetime, as the reference's buffers do.
 */
#ifndef HPTRACING_HPP
#define HPTRACING_HPP

#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "hptracing.h"

namespace hpt {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &msg) : std::runtime_error(msg), code(c) {}
};

inline void check_host(int rc) {
    if (rc != HPT_OK) throw Error(rc, hpt_host_last_error());
}
inline void check(int rc) {
    if (rc != HPT_OK) throw Error(rc, hpt_last_error());
}

struct Float3 {
    float x, y, z;
};

/* hp::KDTree: loads the OBJ/MTL scene and builds the KD tree on the CPU. */
class KDTree {
  public:
    explicit KDTree(const std::string &filename, const std::string &mtl_basepath = "",
                    int n_threads = 0) {
        hpt_host_scene *h = nullptr;
        check_host(hpt_host_load_obj(filename.c_str(), mtl_basepath.c_str(), n_threads, &h));
        h_.reset(h);
    }
    /* getData(): flattened nodes and leaf lists (kdtree.cpp:248-292). */
    std::pair<std::vector<hpt_kd_node>, std::vector<int32_t>> getData() const {
        hpt_scene_desc d = desc();
        return {std::vector<hpt_kd_node>(d.nodes, d.nodes + d.n_nodes),
                std::vector<int32_t>(d.leaf_data, d.leaf_data + d.n_leaf_data)};
    }
    hpt_scene_desc desc() const {
        hpt_scene_desc d;
        check_host(hpt_host_describe(h_.get(), &d));
        return d;
    }

  private:
    struct Free {
        void operator()(hpt_host_scene *p) const { hpt_host_free(p); }
    };
    std::unique_ptr<hpt_host_scene, Free> h_;
};

/* hp::TraceRunner: device-resident scene plus the render call. */
class TraceRunner {
  public:
    explicit TraceRunner(const KDTree &scene) {
        hpt_scene_desc d = scene.desc();
        hpt_scene *s = nullptr;
        check(hpt_scene_create(&d, &s));
        s_.reset(s);
    }
    explicit TraceRunner(const std::unique_ptr<KDTree> &scene) : TraceRunner(*scene) {}

    /* TraceRunner::run (trace_runner.h:53-65): result_data receives
     * sample_x * sample_y RGBA8 pixels. */
    void run(unsigned char *result_data, Float3 view_p, Float3 top_dir, Float3 right_dir,
             float width, float height, int sample_x, int sample_y, int supersample_x,
             int supersample_y, Float3 background_color, int sample = 10, int depth = 6,
             int disable_diffuse = 0, float brightness = 1.0f) {
        hpt_view v = {};
        auto put = [](float *dst, Float3 f) {
            dst[0] = f.x;
            dst[1] = f.y;
            dst[2] = f.z;
        };
        put(v.view, view_p);
        put(v.up, top_dir);
        put(v.right, right_dir);
        put(v.background, background_color);
        v.width = width;
        v.height = height;
        v.sample_x = sample_x;
        v.sample_y = sample_y;
        v.supersample_x = supersample_x;
        v.supersample_y = supersample_y;
        v.sample = sample;
        v.depth = depth;
        v.disable_diffuse = disable_diffuse;
        v.brightness = brightness;
        v.seed = seed_++;
        check(hpt_render(s_.get(), &v, result_data, nullptr, &last_stats));
    }

    void set_seed(uint64_t seed) { seed_ = seed; }
    const hpt_render_stats &stats() const { return last_stats; }
    hpt_scene *handle() const { return s_.get(); }

  private:
    struct Free {
        void operator()(hpt_scene *p) const { hpt_scene_destroy(p); }
    };
    std::unique_ptr<hpt_scene, Free> s_;
    uint64_t seed_ = 1;
    hpt_render_stats last_stats = {};
};

}  // namespace hpt

#endif /* HPTRACING_HPP */
