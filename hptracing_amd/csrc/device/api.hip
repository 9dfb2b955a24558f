// C ABI of the GPU side: scene upload (hpt_scene_*) and the ray-cast
// entry points (hpt_cast*), declared in include/hptracing.h.  Replaces the
// constructor of the OpenCL host layer hp::TraceRunner
// (src/hp/trace_runner.cpp:41-108) and exposes kdtree_intersect alone.
#include <cmath>

#include "internal.hpp"

using namespace hpt;
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
        // escape_indices follows parent links and reads the node they name:
        // every link must name a node (or be -1) before it runs.
        for (int64_t i = 0; i < nn; ++i) {
            const hpt_kd_node &nd = nodes[i];
            if (nd.child < -1 || nd.child >= nn || nd.parent < -1 || nd.parent >= nn ||
                nd.sibling < -1 || nd.sibling >= nn)
                fail(HPT_ERR_SCENE, "KD node link out of range");
        }
        std::vector<int32_t> esc = escape_indices(nodes, nn);
        std::vector<float> pnodes(8 * nn);
        std::vector<float> recs;
        int64_t entries = 0;
        // Box coordinates must be 0 or of magnitude in [2^-60, 2^80] for the
        // reciprocal slab test to be certifiable (box_hit_rcp).
        bool boxes_fast = true;
        auto coord_ok = [](float c) {
            const float a = std::fabs(c);
            return a == 0.f || (a >= 0x1p-60f && a <= 0x1p80f);
        };
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
        std::vector<int32_t> leaf_ids(std::max<int64_t>(entries, 1), 0);
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
                    leaf_ids[rec] = tri;
                    ++rec;
                }
            } else {
                a = 0;  // empty leaf
                b = 0;
            }
            for (int k = 0; k < 3; ++k)
                boxes_fast = boxes_fast && coord_ok(nd.box_start[k]) && coord_ok(nd.box_end[k]);
            float *p = &pnodes[8 * i];
            p[0] = nd.box_start[0]; p[1] = nd.box_start[1]; p[2] = nd.box_start[2];
            std::memcpy(&p[3], &a, 4);
            p[4] = nd.box_end[0]; p[5] = nd.box_end[1]; p[6] = nd.box_end[2];
            std::memcpy(&p[7], &b, 4);
        }

        std::unique_ptr<hpt_scene> s(new hpt_scene());
        hip_check(hipGetDevice(&s->device), "hipGetDevice");
        for (int k = 0; k < 3; ++k) {  // the root's box orders sorted casts
            s->box_lo[k] = nodes[0].box_start[k];
            s->box_hi[k] = nodes[0].box_end[k];
        }
        int64_t total = 0;
        upload(s->nodes, pnodes.data(), pnodes.size(), total);
        upload(s->tris, recs.data(), recs.size(), total);
        // The same records once per triangle, with the leaf lists as record
        // numbers (the indirect fetch form): 48 bytes per triangle plus 4 per
        // leaf entry instead of 48 per leaf entry.  Records are numbered in
        // the order the leaves first list their triangles (leaves in
        // pre-order, entries in list order), so that a leaf's records, and
        // those of the leaves next to it, lie together; each record keeps
        // its triangle's index in its w slot, which is what a hit reports.
        // The walk tests the entries in the same order either way.
        {
            std::vector<int32_t> number(std::max<int64_t>(nt, 1), -1);
            std::vector<int32_t> order;
            order.reserve(nt);
            for (int64_t e = 0; e < entries; ++e) {
                int32_t &k = number[leaf_ids[e]];
                if (k < 0) {
                    k = (int32_t)order.size();
                    order.push_back(leaf_ids[e]);
                }
                leaf_ids[e] = k;
            }
            for (int32_t tri = 0; tri < nt; ++tri)  // triangles no leaf lists
                if (number[tri] < 0) order.push_back(tri);
            std::vector<float> trecs(12 * std::max<int64_t>(nt, 1), 0.f);
            for (int32_t k = 0; k < nt; ++k) {
                const int32_t tri = order[k];
                const int32_t *t = d->mesh + 4 * (int64_t)tri;
                const float *pa = d->points + 4 * (int64_t)t[0];
                const float *pb = d->points + 4 * (int64_t)t[1];
                const float *pc = d->points + 4 * (int64_t)t[2];
                float *r = &trecs[12 * (int64_t)k];
                r[0] = pa[0]; r[1] = pa[1]; r[2] = pa[2];
                std::memcpy(&r[3], &tri, 4);
                r[4] = pb[0] - pa[0]; r[5] = pb[1] - pa[1]; r[6] = pb[2] - pa[2]; r[7] = 0.f;
                r[8] = pc[0] - pa[0]; r[9] = pc[1] - pa[1]; r[10] = pc[2] - pa[2]; r[11] = 0.f;
            }
            upload(s->tri_recs, trecs.data(), trecs.size(), total);
            upload(s->leaf_ids, leaf_ids.data(), leaf_ids.size(), total);
        }
        upload(s->points, d->points, 4 * (size_t)np, total);
        upload(s->normals, d->normals, 4 * (size_t)d->n_normals, total);
        upload(s->texcoords, d->texcoords, 2 * (size_t)d->n_texcoords, total);
        upload(s->mesh, d->mesh, 4 * (size_t)nt, total);
        upload(s->materials, d->materials, (size_t)nm, total);
        upload(s->lights, d->lights, 4 * (size_t)d->n_lights, total);
        upload(s->textures, d->textures, (size_t)d->n_textures * HPT_TEXTURE_SIZE * HPT_TEXTURE_SIZE * 4, total);
        s->device_bytes = total;
        s->n_nodes = nn;
        s->n_entries = entries;
        s->n_triangles = nt;
        s->n_textures = d->n_textures;

        hpt::DevScene &ds = s->dev;
        ds.nodes = s->nodes.as<float4>();
        ds.n_nodes = (int32_t)nn;
        ds.tris = s->tris.as<float4>();
        ds.tri_recs = s->tri_recs.as<float4>();
        ds.leaf_ids = s->leaf_ids.as<int32_t>();
        ds.n_entries = (int32_t)std::max<int64_t>(entries, 1);
        ds.n_recs = (int32_t)std::max<int64_t>(nt, 1);
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
        ds.fast_div = 1;
        ds.boxes_fast = boxes_fast ? 1 : 0;
        ds.traversal = 2;
        // dot/cross as the reference build evaluates them: on this GPU the
        // reference's kernel runs through ROCm's OpenCL device library,
        // which fuses them (HPT_ARITH_ROCM_OPENCL); its hit distances are
        // then reproduced bit for bit.
        ds.arith_ocl = 1;
        ds.refill = 32;
        ds.node_steps = 24;
        // Buffer loads (fetch bit 2048) address by 32-bit offsets: node,
        // leaf-entry and record arrays under 2 GiB.
        ds.fetch = hpt::fetch_fits(3517, ds) ? 3517 : 1469;
        ds.walk_min = 0;

        int cus = 0;
        hip_check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device),
                  "hipDeviceGetAttribute");
        s->n_cus = std::max(1, cus);
        // Six resident blocks of the cast per CU (24 waves) run the 1080p
        // render faster than the eight that fit with the default form (node
        // pairs, a queue of 8 leaves, per-triangle records): fewer rays in
        // flight share the L1 and L2 better.
        s->cast_grid = std::max(1, cus * std::max(1, std::min(6, hpt::occupancy_blocks_cast())));
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

int hpt_scene_get_tuning(const hpt_scene *s, hpt_tuning *out) {
    return guarded([&] {
        if (!s || !out) fail(HPT_ERR_INVALID, "null argument");
        const hpt::LaunchForm f = s->form();
        out->fast_division = f.dev.fast_div;
        out->cast_blocks_per_cu = f.cast_grid / std::max(1, s->n_cus);
        out->shade_blocks_per_cu = f.shade_grid / std::max(1, s->n_cus);
        out->traversal = f.dev.traversal;
        out->refill_threshold = f.dev.refill;
        out->node_steps = f.dev.node_steps;
        out->fetch = f.dev.fetch;
        out->walk_min = f.dev.walk_min;
        out->sort_rays = f.sort_rays;
    });
}

int hpt_scene_set_tuning(hpt_scene *s, const hpt_tuning *t) {
    return guarded([&] {
        if (!s || !t) fail(HPT_ERR_INVALID, "null argument");
        if (t->cast_blocks_per_cu < 1 || t->cast_blocks_per_cu > 64 ||
            t->shade_blocks_per_cu < 1 || t->shade_blocks_per_cu > 64 ||
            t->traversal < 0 || t->traversal > 2 || t->refill_threshold < 1 ||
            t->refill_threshold > 64 || t->node_steps < 1 || t->node_steps > 4096 ||
            !hpt::valid_fetch(t->fetch) || !hpt::fetch_fits(t->fetch, s->dev) ||
            t->walk_min < 0 || t->walk_min > 64 ||
            t->sort_rays < 0 || t->sort_rays > 1)
            fail(HPT_ERR_INVALID, "tuning value out of range");
        std::lock_guard<std::mutex> lock(s->mu);
        std::lock_guard<std::mutex> tlock(s->tune_mu);
        s->dev.fast_div = t->fast_division ? 1 : 0;
        s->dev.traversal = t->traversal;
        s->dev.refill = t->refill_threshold;
        s->dev.node_steps = t->node_steps;
        s->dev.fetch = t->fetch;
        s->dev.walk_min = t->walk_min;
        s->sort_rays = t->sort_rays;
        s->cast_grid = t->cast_blocks_per_cu * s->n_cus;
        s->shade_grid = t->shade_blocks_per_cu * s->n_cus;
    });
}

int hpt_scene_set_arithmetic(hpt_scene *s, int mode) {
    return guarded([&] {
        if (!s) fail(HPT_ERR_INVALID, "null scene");
        if (mode != HPT_ARITH_LITERAL && mode != HPT_ARITH_ROCM_OPENCL)
            fail(HPT_ERR_INVALID, "unknown arithmetic mode");
        std::lock_guard<std::mutex> lock(s->mu);
        std::lock_guard<std::mutex> tlock(s->tune_mu);
        s->dev.arith_ocl = mode == HPT_ARITH_ROCM_OPENCL ? 1 : 0;
    });
}

static int cast_impl(hpt_scene *s, const float *org, const float *dir, int64_t n, int32_t *tri_id,
                     float *t, int32_t *counts, void *stream) {
    return guarded([&] {
        if (!s) fail(HPT_ERR_INVALID, "null scene");
        // The shared ray counter is 32-bit and runs past n by at most two
        // 64-ray reservations per wave.
        if (n < 0 || n > ((int64_t)1 << 30)) fail(HPT_ERR_INVALID, "ray count out of range (> 2^30)");
        if (n == 0) return;
        if (!org || !dir || !tri_id || !t) fail(HPT_ERR_INVALID, "null buffer");
        // NULL is HIP's null stream, so a cast follows whatever the caller
        // enqueued there to produce its rays.
        hipStream_t st = (hipStream_t)stream;
        const hpt::LaunchForm f = s->form();
        if (!counts && f.sort_rays && n >= hpt::kSortMinRays) {
            hpt::cast_sorted(s, f, org, dir, n, tri_id, t, st);
            return;
        }
        hpt::StreamCounter wc(st);
        hpt::launch_cast(f.dev, org, dir, nullptr, nullptr, hpt::CountRef{}, n, wc.p, tri_id, t,
                         counts, nullptr, f.cast_grid, st);
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

}  // extern "C"
