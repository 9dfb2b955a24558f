// C ABI of the CPU side (hpt_host_*, declared in include/hptracing.h).
#include <cstring>
#include <new>
#include <string>

#include "host.h"

struct hpt_host_scene {
    hpt::HostScene s;
};

namespace {
thread_local std::string g_host_error;

int fail(int code, const std::string &msg) {
    g_host_error = msg;
    return code;
}

template <class F>
int guarded(F &&f) {
    try {
        g_host_error.clear();
        return f();
    } catch (const hpt::Error &e) {
        return fail(e.code, e.what());
    } catch (const std::bad_alloc &) {
        return fail(HPT_ERR_NOMEM, "host allocation failed");
    } catch (const std::exception &e) {
        return fail(HPT_ERR_INVALID, e.what());
    }
}
}  // namespace

extern "C" {

const char *hpt_host_last_error(void) { return g_host_error.c_str(); }

int hpt_host_load_obj(const char *obj_path, const char *mtl_basepath, int n_threads,
                      hpt_host_scene **out) {
    if (!obj_path || !out) return fail(HPT_ERR_INVALID, "null argument");
    *out = nullptr;
    return guarded([&] {
        std::unique_ptr<hpt_host_scene> h(new hpt_host_scene());
        hpt::scene_from_obj(obj_path, mtl_basepath ? mtl_basepath : "", h->s);
        hpt::build_light_table(h->s);
        hpt::build_kdtree(h->s, n_threads);
        *out = h.release();
        return HPT_OK;
    });
}

int hpt_host_from_arrays(const float *points_xyz, int64_t n_points, const int32_t *mesh,
                         int64_t n_triangles, const hpt_material *materials,
                         int64_t n_materials, int n_threads, hpt_host_scene **out) {
    if (!points_xyz || !mesh || !materials || !out) return fail(HPT_ERR_INVALID, "null argument");
    if (n_points <= 0 || n_triangles <= 0 || n_materials <= 0 || n_points > INT32_MAX ||
        n_triangles > INT32_MAX)
        return fail(HPT_ERR_INVALID, "array size out of range");
    *out = nullptr;
    return guarded([&] {
        std::unique_ptr<hpt_host_scene> h(new hpt_host_scene());
        hpt::HostScene &s = h->s;
        s.materials.assign(materials, materials + n_materials);
        s.points.resize(4 * n_points);
        for (int64_t i = 0; i < n_points; ++i) {
            s.points[4 * i + 0] = points_xyz[3 * i + 0];
            s.points[4 * i + 1] = points_xyz[3 * i + 1];
            s.points[4 * i + 2] = points_xyz[3 * i + 2];
            s.points[4 * i + 3] = 0.f;
        }
        s.mesh.assign(mesh, mesh + 4 * n_triangles);
        for (int64_t i = 0; i < n_triangles; ++i) {
            for (int k = 0; k < 3; ++k)
                if (s.mesh[4 * i + k] < 0 || s.mesh[4 * i + k] >= n_points)
                    throw hpt::Error(HPT_ERR_SCENE, "triangle references a missing vertex");
            if (s.mesh[4 * i + 3] < 0 || s.mesh[4 * i + 3] >= n_materials)
                throw hpt::Error(HPT_ERR_SCENE, "triangle references a missing material");
        }
        hpt::build_light_table(s);
        hpt::build_kdtree(s, n_threads);
        *out = h.release();
        return HPT_OK;
    });
}

int hpt_host_kd_count(const float *points_xyz, int64_t n_points, const int32_t *mesh,
                      int64_t n_triangles, int n_threads, hpt_kd_counts *out) {
    if (!points_xyz || !mesh || !out) return fail(HPT_ERR_INVALID, "null argument");
    if (n_points <= 0 || n_triangles <= 0 || n_points > INT32_MAX || n_triangles > INT32_MAX)
        return fail(HPT_ERR_INVALID, "array size out of range");
    return guarded([&] {
        std::vector<float> points(4 * n_points, 0.f);
        for (int64_t i = 0; i < n_points; ++i)
            for (int k = 0; k < 3; ++k) points[4 * i + k] = points_xyz[3 * i + k];
        std::vector<int32_t> m(mesh, mesh + 4 * n_triangles);
        for (int64_t i = 0; i < n_triangles; ++i)
            for (int k = 0; k < 3; ++k)
                if (m[4 * i + k] < 0 || m[4 * i + k] >= n_points)
                    throw hpt::Error(HPT_ERR_SCENE, "triangle references a missing vertex");
        const hpt::KdCounts c = hpt::count_kdtree(points, m, n_threads);
        std::memset(out, 0, sizeof(*out));
        out->n_nodes = c.nodes;
        out->n_leaves = c.leaves;
        out->n_leaf_entries = c.entries;
        out->n_leaf_ints = c.leaves + c.entries;
        out->max_leaf = c.max_leaf;
        out->max_depth = c.max_depth;
        out->seconds = c.seconds;
        return HPT_OK;
    });
}

int hpt_host_read_image(const char *path, int32_t *w, int32_t *h, int32_t *c, float *out,
                        int64_t capacity) {
    if (!path || !w || !h || !c) return fail(HPT_ERR_INVALID, "null argument");
    return guarded([&] {
        const hpt::Image img = hpt::load_texture(path);
        *w = img.w;
        *h = img.h;
        *c = img.c;
        if (out) {
            if (capacity < (int64_t)img.data.size())
                throw hpt::Error(HPT_ERR_INVALID, "image buffer too small");
            std::memcpy(out, img.data.data(), img.data.size() * sizeof(float));
        }
        return HPT_OK;
    });
}

int hpt_host_describe(const hpt_host_scene *scene, hpt_scene_desc *out) {
    if (!scene || !out) return fail(HPT_ERR_INVALID, "null argument");
    const hpt::HostScene &s = scene->s;
    std::memset(out, 0, sizeof(*out));
    out->points = s.points.data();
    out->n_points = s.n_points();
    out->normals = s.normals.data();
    out->n_normals = (int64_t)s.normals.size() / 4;
    out->texcoords = s.texcoords.data();
    out->n_texcoords = (int64_t)s.texcoords.size() / 2;
    out->mesh = s.mesh.data();
    out->n_triangles = s.n_triangles();
    out->materials = s.materials.data();
    out->n_materials = (int64_t)s.materials.size();
    out->lights = s.lights.data();
    out->n_lights = (int64_t)s.lights.size() / 4;
    out->nodes = s.nodes.data();
    out->n_nodes = (int64_t)s.nodes.size();
    out->leaf_data = s.leaf_data.data();
    out->n_leaf_data = (int64_t)s.leaf_data.size();
    out->textures = s.texture_data.data();
    out->n_textures = (int64_t)s.texture_names.size();
    return HPT_OK;
}

int hpt_host_material_from_mtl(const float ka[3], const float kd[3], const float ks[3],
                               const float ke[3], float ior, float dissolve, float shininess,
                               hpt_material *out) {
    if (!ka || !kd || !ks || !ke || !out) return fail(HPT_ERR_INVALID, "null argument");
    *out = hpt::material_from_mtl(ka, kd, ks, ke, ior, dissolve, shininess);
    return HPT_OK;
}

void hpt_host_free(hpt_host_scene *scene) { delete scene; }

double hpt_host_build_seconds(const hpt_host_scene *scene) {
    return scene ? scene->s.build_seconds : 0.0;
}

}  // extern "C"
