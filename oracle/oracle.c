/*
 * CPU oracle of hpTracing's path-tracing hot path -- TEST INFRASTRUCTURE.
 * See oracle.h for the conventions and for what pins each function.
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef struct { float x, y, z; } v3;

static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 ld3(const float *p) { return mk(p[0], p[1], p[2]); }
static inline v3 ldf3(orc_f3 p) { return mk(p.x, p.y, p.z); }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mulv(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 scl(float s, v3 a) { return mk(s * a.x, s * a.y, s * a.z); }
static inline v3 neg(v3 a) { return mk(-a.x, -a.y, -a.z); }
static inline float dot(v3 a, v3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
static inline v3 cross(v3 a, v3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline float length3(v3 a) { return sqrtf(dot(a, a)); }
static inline v3 normalize3(v3 a) {
    float l = length3(a);
    return mk(a.x / l, a.y / l, a.z / l);
}

#define PI_F 3.1415926535f
#define GENERAL_THRESHOLD 1e-4f
#define STEP_EPSILON 1e-5f

static int threads_or_all(int n) {
#ifdef _OPENMP
    return n > 0 ? n : omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* ------------------------------------------------------------ kernel.cl:51-83 */
int orc_box_intersect(const float bs[3], const float be[3], const float sp[3], const float d[3]) {
    if (sp[0] >= bs[0] && sp[0] <= be[0] && sp[1] >= bs[1] && sp[1] <= be[1] && sp[2] >= bs[2] &&
        sp[2] <= be[2])
        return 1;
    float t0, t1, mins[3], maxs[3];
    for (int k = 0; k < 3; ++k) {
        t0 = (bs[k] - sp[k]) / d[k];
        t1 = (be[k] - sp[k]) / d[k];
        mins[k] = fminf(t0, t1);
        maxs[k] = fmaxf(t0, t1);
    }
    float max_of_mins = fmaxf(fmaxf(mins[0], mins[1]), mins[2]);
    float min_of_maxs = fminf(fminf(maxs[0], maxs[1]), maxs[2]);
    if (min_of_maxs <= 0) return 0;
    return max_of_mins <= min_of_maxs;
}

/* ------------------------------------------------------------ kernel.cl:85-109 */
/* dot/cross of the intersection test: ORC_ARITH_LITERAL rounds every product
 * and sum; ORC_ARITH_ROCM_OPENCL fuses them as ROCm's OpenCL device library
 * does (tests/golden/rocm_ocl_builtins/samples.npz holds the library's
 * outputs). */
static int g_arith = ORC_ARITH_ROCM_OPENCL;
void orc_set_arithmetic(int mode) { g_arith = mode; }

static inline float dot_ocl(v3 a, v3 b) { return fmaf(a.z, b.z, fmaf(a.y, b.y, a.x * b.x)); }
static inline v3 cross_ocl(v3 a, v3 b) {
    return mk(fmaf(a.y, b.z, -(a.z * b.y)), fmaf(a.z, b.x, -(a.x * b.z)),
              fmaf(a.x, b.y, -(a.y * b.x)));
}

static inline float single_intersect(v3 sp, v3 d, v3 pa, v3 pb, v3 pc) {
    const int ocl = g_arith == ORC_ARITH_ROCM_OPENCL;
    v3 e1 = sub(pb, pa), e2 = sub(pc, pa);
    v3 P = ocl ? cross_ocl(d, e2) : cross(d, e2);
    float det = ocl ? dot_ocl(e1, P) : dot(e1, P);
    /* `det > EPSILON && det < EPSILON` (kernel.cl:93) can never hold. */
    float inv_det = 1.f / det;
    v3 T = sub(sp, pa);
    float u = (ocl ? dot_ocl(T, P) : dot(T, P)) * inv_det;
    if (u < 0.f || u > 1.f) return -1;
    v3 Q = ocl ? cross_ocl(T, e1) : cross(T, e1);
    float v = (ocl ? dot_ocl(d, Q) : dot(d, Q)) * inv_det;
    if (v < 0.f || u + v > 1.f) return -1;
    float t = (ocl ? dot_ocl(e2, Q) : dot(e2, Q)) * inv_det;
    if (t > 0) return t;
    return -1;
}

void orc_dot_cross(int mode, const float *a, const float *b, int64_t n, float *out) {
    for (int64_t i = 0; i < n; ++i) {
        v3 x = ld3(a + 3 * i), y = ld3(b + 3 * i);
        float d = mode == ORC_ARITH_ROCM_OPENCL ? dot_ocl(x, y) : dot(x, y);
        v3 c = mode == ORC_ARITH_ROCM_OPENCL ? cross_ocl(x, y) : cross(x, y);
        out[4 * i] = d;
        out[4 * i + 1] = c.x;
        out[4 * i + 2] = c.y;
        out[4 * i + 3] = c.z;
    }
}

float orc_single_intersect(const float sp[3], const float d[3], const float pa[3],
                           const float pb[3], const float pc[3]) {
    return single_intersect(ld3(sp), ld3(d), ld3(pa), ld3(pb), ld3(pc));
}

/* ----------------------------------------------------------- kernel.cl:111-185 */
static int32_t kd_search(const orc_scene *s, v3 sp, v3 d, float *t_out, int32_t *counts) {
    int match[ORC_MAX_MATCH_DATA];
    int nmatch = 0, visits = 0, tested = 0, dropped = 0, visits_cap = -1;
    int node_index = 0, come_from_child = 0;
    const float spa[3] = {sp.x, sp.y, sp.z}, da[3] = {d.x, d.y, d.z};
    for (;;) {
        int goto_child = 0;
        const orc_node *node = &s->nodes[node_index];
        if (come_from_child == 0) {
            ++visits;
            const float bs[3] = {node->box_start.x, node->box_start.y, node->box_start.z};
            const float be[3] = {node->box_end.x, node->box_end.y, node->box_end.z};
            if (orc_box_intersect(bs, be, spa, da)) {
                if (node->child < 0) {
                    if (node->data >= 0) {
                        if (nmatch < ORC_MAX_MATCH_DATA) {
                            match[nmatch++] = node->data;
                            if (nmatch == ORC_MAX_MATCH_DATA) visits_cap = visits;
                        } else {
                            ++dropped;
                        }
                    }
                } else {
                    goto_child = 1;
                }
            }
        }
        if (goto_child) {
            come_from_child = 0;
            node_index = node->child;
        } else if (node->sibling >= 0) {
            come_from_child = 0;
            node_index = node->sibling;
        } else if (node->parent >= 0) {
            come_from_child = 1;
            node_index = node->parent;
        } else {
            break;
        }
    }
    int32_t geo_id = -1;
    float best = -42;
    for (int i = 0; i < nmatch; ++i) {
        const int32_t *data = s->leaf_data + match[i];
        int n = data[0];
        for (int x = 0; x < n; ++x) {
            int32_t tri = data[x + 1];
            const int32_t *g = &s->mesh[4 * (int64_t)tri];
            float r = single_intersect(sp, d, ldf3(s->points[g[0]]), ldf3(s->points[g[1]]),
                                       ldf3(s->points[g[2]]));
            ++tested;
            if (r > 0 && (best < 0 || r < best)) {
                best = r;
                geo_id = tri;
            }
        }
    }
    if (counts) {
        counts[0] = visits_cap >= 0 ? visits_cap : visits;
        counts[1] = nmatch;
        counts[2] = tested;
        counts[3] = dropped;
        counts[4] = visits;
    }
    *t_out = best;
    return geo_id;
}

void orc_kd_cast(const orc_scene *s, const float *org, const float *dir, int64_t n,
                 int32_t *tri_id, float *t, int32_t *counts, int n_threads) {
    int nt = threads_or_all(n_threads);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
    for (int64_t i = 0; i < n; ++i) {
        float tt;
        tri_id[i] = kd_search(s, ld3(org + 3 * i), ld3(dir + 3 * i), &tt, counts ? counts + ORC_CAST_COUNTS * i : NULL);
        t[i] = tt;
    }
}

/* ----------------------------------------------------------- kernel.cl:210-245 */
void orc_naive_cast(const orc_scene *s, const float *org, const float *dir, int64_t n,
                    int32_t *tri_id, float *t, int n_threads) {
    int nt = threads_or_all(n_threads);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (int64_t i = 0; i < n; ++i) {
        v3 sp = ld3(org + 3 * i), d = ld3(dir + 3 * i);
        int32_t geo_id = -1;
        float best = -42;
        for (int64_t k = 0; k < s->n_triangles; ++k) {
            const int32_t *g = &s->mesh[4 * k];
            float r = single_intersect(sp, d, ldf3(s->points[g[0]]), ldf3(s->points[g[1]]),
                                       ldf3(s->points[g[2]]));
            if (r > 0 && (best < 0 || r < best)) {
                best = r;
                geo_id = (int32_t)k;
            }
        }
        tri_id[i] = geo_id;
        t[i] = best;
    }
}

/* ------------------------------------------------------------ kernel.cl:26-47 */
int64_t orc_rand(int64_t *seed) {
    uint64_t s = (uint64_t)*seed;
    s = (s * 0x5DEECE66DULL + 0xBULL) & ((1ULL << 48) - 1);
    *seed = (int64_t)s;
    return (int64_t)((s >> 16) & 0xFFFFFFFFULL);
}

float orc_randf(int64_t *seed) { return (float)orc_rand(seed) / (float)0xFFFFFFFFLL; }

static inline v3 randf3(int64_t *seed) {
    float alpha = orc_randf(seed) * PI_F;
    float beta = orc_randf(seed) * PI_F;
    float cos_beta = cosf(beta);
    return mk(cos_beta * cosf(alpha), cos_beta * sinf(alpha), sinf(beta));
}

void orc_randf3(int64_t *seed, float out[3]) {
    v3 r = randf3(seed);
    out[0] = r.x;
    out[1] = r.y;
    out[2] = r.z;
}

static inline uint64_t splitmix(uint64_t z) {
    z += 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

uint64_t orc_path_seed(uint64_t seed, int64_t ray, int32_t pass) {
    return splitmix(splitmix(seed ^ splitmix((uint64_t)ray)) + (uint64_t)(uint32_t)pass) &
           ((1ULL << 48) - 1);
}

/* ----------------------------------------------------------- kernel.cl:559-613 */
/* Direction of the camera ray (gx, gy) of set_viewdirs / set_viewdirs_vr. */
static inline v3 pinhole_dir(v3 up, v3 right, float width, float height, int sx, int sy, int gx,
                             int gy) {
    v3 nd = cross(up, right);
    nd = add(nd, scl((float)gy / (float)sy * height - height * 0.5f, up));
    nd = add(nd, scl((float)gx / (float)sx * width - width * 0.5f, right));
    return normalize3(nd);
}

static inline v3 vr_dir(int x_shift, int sample_y, int gx, int gy) {
    float lon = (float)(gx + x_shift) / (float)(sample_y * 2) * PI_F * 2.f - PI_F;
    float lat = (float)gy / (float)sample_y * PI_F - PI_F / 2.0f;
    v3 d = mk(cosf(lat) * cosf(lon), sinf(lat), cosf(lat) * sinf(lon));
    return normalize3(d);
}

void orc_set_viewdirs(const float view[3], const float up_[3], const float right_[3], float width,
                      float height, int sx, int sy, float *dirs) {
    (void)view;
    v3 up = ld3(up_), right = ld3(right_);
    for (int gy = 0; gy < sy; ++gy)
        for (int gx = 0; gx < sx; ++gx) {
            v3 r = pinhole_dir(up, right, width, height, sx, sy, gx, gy);
            int64_t idx = (int64_t)sx * gy + gx;
            dirs[3 * idx + 0] = r.x;
            dirs[3 * idx + 1] = r.y;
            dirs[3 * idx + 2] = r.z;
        }
}

void orc_set_viewdirs_vr(int x_shift, int sample_x, int sample_y, int grid_x, int grid_y,
                         float *dirs) {
    for (int gy = 0; gy < grid_y; ++gy)
        for (int gx = 0; gx < grid_x; ++gx) {
            v3 r = vr_dir(x_shift, sample_y, gx, gy);
            int64_t idx = (int64_t)sample_x * gy + gx;
            dirs[3 * idx + 0] = r.x;
            dirs[3 * idx + 1] = r.y;
            dirs[3 * idx + 2] = r.z;
        }
}

/* ----------------------------------------------------------- kernel.cl:615-652 */
static inline uint8_t unorm8(float f) {
    /* convert_uchar_sat_rte(f * 255.0f); NaN converts to 0. */
    float x = f * 255.0f;
    if (!(x > 0.f)) return 0;
    if (x >= 255.f) return 255;
    return (uint8_t)nearbyintf(x);
}

void orc_get_image(const float *accum, int width, int height, int ssx, int ssy, int spp,
                   float brightness, uint8_t *rgba) {
    for (int gy = 0; gy < height; ++gy)
        for (int gx = 0; gx < width; ++gx) {
            float px = 0, py = 0, pz = 0;
            for (int i = 0; i < ssx; ++i)
                for (int j = 0; j < ssy; ++j) {
                    int64_t row = (int64_t)gy * ssy + j, col = (int64_t)gx * ssx + i;
                    int64_t index = row * width * ssx + col;
                    px += accum[index * 3];
                    py += accum[index * 3 + 1];
                    pz += accum[index * 3 + 2];
                }
            float n = (float)(ssx * ssy), sp = (float)spp;
            px /= n; py /= n; pz /= n;
            px /= sp; py /= sp; pz /= sp;
            px *= brightness; py *= brightness; pz *= brightness;
            px = 1.f - expf(-px);
            py = 1.f - expf(-py);
            pz = 1.f - expf(-pz);
            uint8_t *o = rgba + 4 * ((int64_t)gy * width + gx);
            o[0] = unorm8(px);
            o[1] = unorm8(py);
            o[2] = unorm8(pz);
            o[3] = 0;
        }
}

/* --------------------------- read_imagef, OpenCL 1.2 section 8.2 (REPEAT, LINEAR) */
static inline v3 texel(const orc_scene *s, int layer, int i, int j) {
    const uint8_t *p = s->textures + (((int64_t)layer * 512 + j) * 512 + i) * 4;
    return mk(p[0] / 255.f, p[1] / 255.f, p[2] / 255.f);
}

void orc_read_texture(const orc_scene *s, int layer, float su, float sv, float out[3]) {
    const float W = 512.f, H = 512.f;
    if (layer < 0) layer = 0;
    if (layer > s->n_textures - 1) layer = (int)s->n_textures - 1;
    float u = (su - floorf(su)) * W;
    float v = (sv - floorf(sv)) * H;
    float fu = floorf(u - 0.5f), fv = floorf(v - 0.5f);
    int i0 = (int)fu, j0 = (int)fv;
    int i1 = i0 + 1, j1 = j0 + 1;
    if (i0 < 0) i0 = 512 + i0;
    if (j0 < 0) j0 = 512 + j0;
    if (i1 > 511) i1 = i1 - 512;
    if (j1 > 511) j1 = j1 - 512;
    /* NaN coordinates give out-of-range indices: fetch texel 0 like a
     * clamped address would (parity unpinned for NaN texture coordinates). */
    if (i0 < 0 || i0 > 511 || i1 < 0 || i1 > 511) i0 = i1 = 0;
    if (j0 < 0 || j0 > 511 || j1 < 0 || j1 > 511) j0 = j1 = 0;
    float a = (u - 0.5f) - fu, b = (v - 0.5f) - fv;
    v3 t00 = texel(s, layer, i0, j0), t10 = texel(s, layer, i1, j0);
    v3 t01 = texel(s, layer, i0, j1), t11 = texel(s, layer, i1, j1);
    float w00 = (1 - a) * (1 - b), w10 = a * (1 - b), w01 = (1 - a) * b, w11 = a * b;
    out[0] = ((w00 * t00.x + w10 * t10.x) + w01 * t01.x) + w11 * t11.x;
    out[1] = ((w00 * t00.y + w10 * t10.y) + w01 * t01.y) + w11 * t11.y;
    out[2] = ((w00 * t00.z + w10 * t10.z) + w01 * t01.z) + w11 * t11.z;
}

/* ------------------------------------------------------ per-ray record (unit_data) */
typedef struct {
    int32_t orig_id;
    v3 strength, start_p, in_dir;
    int32_t geometry[4];
    float intersect_number;
    v3 intersect_p, normal;
    float optical_density, shininess;
} unit;

enum { BR_NONE = -1, BR_REFRACT = 0, BR_SPECULAR = 1, BR_DIFFUSE = 2, BR_LIGHT = 3 };

/* s1_run body (kernel.cl:283-408) for one record.  Returns the branch the
 * ray was appended to, or BR_NONE if it was dropped. */
static int shade(const orc_scene *s, unit *u, float *accum, int64_t *seed, int disable_diffuse) {
    const int32_t *g = u->geometry;
    v3 A = ldf3(s->points[g[0]]), B = ldf3(s->points[g[1]]), Cc = ldf3(s->points[g[2]]);
    const orc_material *mat = &s->materials[g[3]];
    v3 P = add(u->start_p, scl(u->intersect_number, u->in_dir));
    v3 f1 = sub(A, P), f2 = sub(B, P), f3 = sub(Cc, P);
    float area_all = length3(cross(sub(B, A), sub(Cc, A)));
    float a1 = length3(cross(f2, f3)) / area_all;
    float a2 = length3(cross(f1, f3)) / area_all;
    float a3 = length3(cross(f1, f2)) / area_all;
    v3 normal;
    if (s->n_normals == 0) {
        normal = normalize3(cross(sub(B, A), sub(Cc, A)));
    } else {
        normal = add(add(scl(a1, ldf3(s->normals[g[0]])), scl(a2, ldf3(s->normals[g[1]]))),
                     scl(a3, ldf3(s->normals[g[2]])));
    }
    float uvx = 0, uvy = 0;
    if (mat->ambient_texture_id >= 0 || mat->diffuse_texture_id >= 0) {
        if (s->n_texcoords == 0) {
            float l = sqrtf(P.x * P.x + P.y * P.y);
            uvx = P.x / l;
            uvy = P.y / l;
        } else {
            const float *ta = s->texcoords + 2 * (int64_t)g[0];
            const float *tb = s->texcoords + 2 * (int64_t)g[1];
            const float *tc = s->texcoords + 2 * (int64_t)g[2];
            uvx = (a1 * ta[0] + a2 * tb[0]) + a3 * tc[0];
            uvy = (a1 * ta[1] + a2 * tb[1]) + a3 * tc[1];
        }
        uvy = 1.f - uvy;
    }
    v3 ambient = ldf3(mat->ambient);
    if (mat->ambient_texture_id >= 0) {
        float c[3];
        orc_read_texture(s, mat->ambient_texture_id, uvx, uvy, c);
        ambient = ld3(c);
    }
    v3 result = mulv(u->strength, add(ambient, ldf3(mat->emission)));
    result = scl(fabsf(dot(normal, u->in_dir)), result);
    float *target = accum + 3 * (int64_t)u->orig_id;
    target[0] += result.x;
    target[1] += result.y;
    target[2] += result.z;

    if (length3(u->strength) < GENERAL_THRESHOLD) return BR_NONE;
    u->intersect_p = P;
    u->normal = normal;
    float r = orc_randf(seed);
    if (r < mat->specular_possibility) {
        u->strength = mulv(u->strength, ldf3(mat->specular));
        u->shininess = mat->shininess;
        return BR_SPECULAR;
    }
    if (r < mat->refract_possibility) {
        u->strength = scl(1.0f - mat->dissolve, u->strength);
        u->optical_density = mat->optical_density;
        return BR_REFRACT;
    }
    v3 diffuse = ldf3(mat->diffuse);
    if (mat->diffuse_texture_id >= 0) {
        float c[3];
        orc_read_texture(s, mat->diffuse_texture_id, uvx, uvy, c);
        diffuse = ld3(c);
    }
    if (r < mat->diffuse_possibility && disable_diffuse == 0) {
        u->strength = mulv(u->strength, diffuse);
        return BR_DIFFUSE;
    }
    u->strength = mulv(u->strength, diffuse);
    return BR_LIGHT;
}

/* s2_refract_run (kernel.cl:413-449). */
static void bounce_refract(unit *u) {
    float cos_alpha = dot(u->in_dir, neg(u->normal));
    float reverse = 1.0f;
    if (cos_alpha < 0) reverse = -1;
    float alpha = acosf(cos_alpha);
    v3 p = scl(cos_alpha, u->normal);
    v3 q = normalize3(add(u->in_dir, p));
    float sin_beta = sinf(alpha) * powf(u->optical_density, -reverse);
    v3 fd;
    if (sin_beta <= 1.f) {
        float beta = asinf(sin_beta);
        fd = add(scl(-reverse * cosf(beta), u->normal), scl(sin_beta, q));
    } else {
        fd = add(add(u->in_dir, p), p);
    }
    u->start_p = add(u->intersect_p, scl(STEP_EPSILON, fd));
    u->in_dir = fd;
}

/* s2_specular_run (kernel.cl:451-478). */
static void bounce_specular(unit *u, int64_t *seed) {
    v3 rv = randf3(seed);
    rv = mk(rv.x / 2.0f / u->shininess, rv.y / 2.0f / u->shininess, rv.z / 2.0f / u->shininess);
    float d = dot(u->in_dir, u->normal);
    v3 projection = scl(d, u->normal);
    v3 refl = sub(u->in_dir, scl(2.0f, projection));
    refl = add(refl, rv);
    u->start_p = add(u->intersect_p, scl(STEP_EPSILON, refl));
    u->in_dir = refl;
}

/* s2_diffuse_run (kernel.cl:480-510). */
static void bounce_diffuse(unit *u, int64_t *seed) {
    int front = dot(u->in_dir, u->normal) < 0;
    v3 p = randf3(seed);
    float dn = dot(p, u->normal);
    if (dn < 0) {
        dn = -dn;
        if (front) p = neg(p);
    }
    u->strength = scl(dn, u->strength);
    u->start_p = add(u->intersect_p, scl(STEP_EPSILON, p));
    u->in_dir = p;
}

/* s2_light_run (kernel.cl:512-557).  Returns 1 if the ray continues. */
static int bounce_light(const orc_scene *s, unit *u, int64_t *seed) {
    if (s->n_lights == 0) return 0;
    int front = dot(u->in_dir, u->normal) < 0;
    int idx = (int)(orc_rand(seed) % s->n_lights);
    const int32_t *l = &s->lights[4 * (int64_t)idx];
    v3 pa = ldf3(s->points[l[0]]), pb = ldf3(s->points[l[1]]), pc = ldf3(s->points[l[2]]);
    float rx = orc_randf(seed), ry = orc_randf(seed);
    if (rx + ry > 1) {
        rx = 1 - rx;
        ry = 1 - ry;
    }
    v3 point = add(add(pa, scl(rx, sub(pb, pa))), scl(ry, sub(pc, pa)));
    v3 p = normalize3(sub(point, u->intersect_p));
    float d = dot(p, u->normal);
    if ((d > 0) == front) {
        u->strength = scl(fabsf(d), u->strength);
        u->start_p = add(u->intersect_p, scl(STEP_EPSILON, p));
        u->in_dir = p;
        return 1;
    }
    return 0;
}

/* kdtree_intersect's write-back (kernel.cl:187-207). Returns 1 on a hit. */
static int intersect_record(const orc_scene *s, unit *u, float *accum, const float bg[3],
                            int sample_n, int32_t *counts) {
    float t;
    int32_t id = kd_search(s, u->start_p, u->in_dir, &t, counts);
    if (id != -1) {
        memcpy(u->geometry, &s->mesh[4 * (int64_t)id], sizeof(u->geometry));
        u->intersect_number = t;
        return 1;
    }
    float *target = accum + 3 * (int64_t)u->orig_id;
    v3 val = scl((float)sample_n, mulv(ld3(bg), u->strength));
    if (sample_n == 1) {
        target[0] += val.x;
        target[1] += val.y;
        target[2] += val.z;
    } else {
        target[0] = val.x;
        target[1] = val.y;
        target[2] = val.z;
    }
    return 0;
}

static int make_camera(const orc_view *v, unit **records_out, int64_t *R_out) {
    const int sx = v->sample_x * v->supersample_x, sy = v->sample_y * v->supersample_y;
    const int64_t R = (int64_t)sx * sy;
    float *dirs = (float *)malloc(sizeof(float) * 3 * R);
    unit *rec = (unit *)calloc(R, sizeof(unit));
    if (!dirs || !rec) {
        free(dirs);
        free(rec);
        return -1;
    }
    if (v->height < 0) {
        int shift = (int)v->right[0];
        orc_set_viewdirs_vr(shift * v->supersample_x, v->sample_x * v->supersample_y,
                            v->sample_y * v->supersample_y, sx, sy, dirs);
    } else {
        orc_set_viewdirs(v->view, v->up, v->right, v->width, v->height, sx, sy, dirs);
    }
    for (int64_t r = 0; r < R; ++r) {
        rec[r].orig_id = (int32_t)r;
        rec[r].strength = mk(1, 1, 1);
        rec[r].start_p = ld3(v->view);
        rec[r].in_dir = ld3(dirs + 3 * r);
    }
    free(dirs);
    *records_out = rec;
    *R_out = R;
    return 0;
}

/* Literal wavefront of TraceRunner::run with lists and a size protocol. */
static int render_wavefront(const orc_scene *s, const orc_view *v, int seed_mode, float *accum,
                            int64_t *live) {
    unit *init = NULL;
    int64_t R = 0;
    if (make_camera(v, &init, &R)) return -1;
    unit *data = (unit *)malloc(sizeof(unit) * R);
    int32_t *lists[6], *cache = (int32_t *)malloc(sizeof(int32_t) * R);
    int64_t *vseed = (int64_t *)malloc(sizeof(int64_t) * R);
    int ok = data && cache && vseed;
    for (int k = 0; k < 6; ++k) {
        lists[k] = (int32_t *)malloc(sizeof(int32_t) * R);
        ok = ok && lists[k];
    }
    if (!ok) return -1;
    memset(accum, 0, sizeof(float) * 3 * R);
    if (live) memset(live, 0, sizeof(int64_t) * v->depth);
    /* Initial seeds: std::rand()-like 31-bit values (trace_runner.cpp:175-178). */
    uint64_t g = v->seed;
    for (int64_t r = 0; r < R; ++r) {
        g = splitmix(g);
        vseed[r] = (int64_t)(g >> 33);
    }
    /* Primary intersect, sample_n = spp, results cached (:192-213). */
    int64_t n1 = 0;
    for (int64_t r = 0; r < R; ++r)
        if (intersect_record(s, &init[r], accum, v->background, v->sample, NULL)) cache[n1++] = (int32_t)r;
    if (live) live[0] += R;
    int64_t sizes[6];
    for (int ii = 0; ii < v->sample; ++ii) {
        memset(sizes, 0, sizeof(sizes));
        sizes[1] = n1;
        memcpy(data, init, sizeof(unit) * R);
        memcpy(lists[1], cache, sizeof(int32_t) * n1);
        if (seed_mode == 0)
            for (int64_t r = 0; r < R; ++r) vseed[r] = (int64_t)orc_path_seed(v->seed, r, ii);
        for (int i = 0; i < v->depth; ++i) {
            if (sizes[0]) {
                if (live) live[i] += sizes[0];
                for (int64_t k = 0; k < sizes[0]; ++k) {
                    int32_t id = lists[0][k];
                    if (intersect_record(s, &data[id], accum, v->background, 1, NULL))
                        lists[1][sizes[1]++] = id;
                }
            }
            sizes[0] = 0;
            for (int64_t k = 0; k < sizes[1]; ++k) {
                int32_t id = lists[1][k];
                int64_t *sd;
                int64_t local;
                if (seed_mode == 1) {
                    local = vseed[k] + k;
                    sd = &local;
                } else {
                    sd = &vseed[id];
                }
                int br = shade(s, &data[id], accum, sd, v->disable_diffuse);
                if (seed_mode == 1 && (br != BR_NONE)) vseed[k] = local;
                if (br != BR_NONE) lists[2 + br][sizes[2 + br]++] = id;
            }
            sizes[1] = 0;
            if (i == v->depth - 1) break;
            for (int b = 0; b < 4; ++b) {
                for (int64_t k = 0; k < sizes[2 + b]; ++k) {
                    int32_t id = lists[2 + b][k];
                    int64_t local = 0, *sd;
                    if (seed_mode == 1) {
                        local = vseed[k] + k;
                        sd = &local;
                    } else {
                        sd = &vseed[id];
                    }
                    int cont = 1;
                    if (b == BR_REFRACT) bounce_refract(&data[id]);
                    else if (b == BR_SPECULAR) bounce_specular(&data[id], sd);
                    else if (b == BR_DIFFUSE) bounce_diffuse(&data[id], sd);
                    else {
                        if (s->n_lights == 0) continue;
                        cont = bounce_light(s, &data[id], sd);
                    }
                    if (seed_mode == 1 && b != BR_REFRACT) vseed[k] = local;
                    if (cont) lists[0][sizes[0]++] = id;
                }
                sizes[2 + b] = 0;
            }
        }
    }
    free(init);
    free(data);
    free(cache);
    free(vseed);
    for (int k = 0; k < 6; ++k) free(lists[k]);
    return 0;
}

/* One camera ray's passes depth-first with per-path seeds (seed_mode 0):
 * the same arithmetic as the wavefront, reordered per path.  Returns rays
 * cast excluding the primary; per-depth counts go to live. */
static int64_t trace_paths(const orc_scene *s, const orc_view *v, const unit *init, int32_t hit_id,
                           float hit_t, float *accum, int64_t *live) {
    int64_t cast = 0;
    for (int ii = 0; ii < v->sample; ++ii) {
        unit u = *init;
        memcpy(u.geometry, &s->mesh[4 * (int64_t)hit_id], sizeof(u.geometry));
        u.intersect_number = hit_t;
        int64_t seed = (int64_t)orc_path_seed(v->seed, init->orig_id, ii);
        for (int i = 0; i < v->depth; ++i) {
            if (i > 0) {
                ++cast;
                if (live) ++live[i];
                if (!intersect_record(s, &u, accum, v->background, 1, NULL)) break;
            }
            int br = shade(s, &u, accum, &seed, v->disable_diffuse);
            if (br == BR_NONE || i == v->depth - 1) break;
            if (br == BR_REFRACT) bounce_refract(&u);
            else if (br == BR_SPECULAR) bounce_specular(&u, &seed);
            else if (br == BR_DIFFUSE) bounce_diffuse(&u, &seed);
            else if (!bounce_light(s, &u, &seed)) break;
        }
    }
    return cast;
}

int64_t orc_render_partial(const orc_scene *s, const orc_view *v, int64_t ray_begin,
                           int64_t ray_end, int n_threads, float *accum) {
    unit *init = NULL;
    int64_t R = 0;
    if (make_camera(v, &init, &R)) return -1;
    if (ray_begin < 0) ray_begin = 0;
    if (ray_end > R) ray_end = R;
    int64_t total = 0;
    int nt = threads_or_all(n_threads);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt) reduction(+ : total)
    for (int64_t r = ray_begin; r < ray_end; ++r) {
        unit u = init[r];
        float t;
        int32_t id = kd_search(s, u.start_p, u.in_dir, &t, NULL);
        total += 1;
        if (id < 0) {
            intersect_record(s, &u, accum, v->background, v->sample, NULL);
            continue;
        }
        total += trace_paths(s, v, &init[r], id, t, accum, NULL);
    }
    free(init);
    return total;
}

int64_t orc_render_rays(const orc_scene *s, const orc_view *v, const int64_t *rays, int64_t n,
                        int n_threads, float *accum) {
    const int sx = v->sample_x * v->supersample_x, sy = v->sample_y * v->supersample_y;
    const int64_t R = (int64_t)sx * sy;
    if (v->height < 0 && v->supersample_x != v->supersample_y) return -1;
    const v3 up = ld3(v->up), right = ld3(v->right);
    const int shift = (int)v->right[0] * v->supersample_x;
    int64_t total = 0;
    int nt = threads_or_all(n_threads);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt) reduction(+ : total)
    for (int64_t k = 0; k < n; ++k) {
        const int64_t r = rays[k];
        if (r < 0 || r >= R) continue;
        const int gx = (int)(r % sx), gy = (int)(r / sx);
        unit u;
        memset(&u, 0, sizeof(u));
        u.orig_id = (int32_t)r;
        u.strength = mk(1, 1, 1);
        u.start_p = ld3(v->view);
        u.in_dir = v->height < 0 ? vr_dir(shift, sy, gx, gy)
                                 : pinhole_dir(up, right, v->width, v->height, sx, sy, gx, gy);
        float t;
        int32_t id = kd_search(s, u.start_p, u.in_dir, &t, NULL);
        total += 1;
        if (id < 0) {
            intersect_record(s, &u, accum, v->background, v->sample, NULL);
            continue;
        }
        total += trace_paths(s, v, &u, id, t, accum, NULL);
    }
    return total;
}

int orc_render(const orc_scene *s, const orc_view *v, int seed_mode, int n_threads, float *accum,
               uint8_t *rgba, int64_t *live) {
    if (v->depth < 1 || v->sample < 1 || v->sample_x < 1 || v->sample_y < 1) return -1;
    if (v->height < 0 && v->supersample_x != v->supersample_y) return -1;
    const int64_t R = (int64_t)v->sample_x * v->supersample_x * v->sample_y * v->supersample_y;
    if (seed_mode == 1) {
        if (render_wavefront(s, v, 1, accum, live)) return -1;
    } else {
        memset(accum, 0, sizeof(float) * 3 * R);
        if (live) memset(live, 0, sizeof(int64_t) * v->depth);
        unit *init = NULL;
        int64_t RR = 0;
        if (make_camera(v, &init, &RR)) return -1;
        int nt = threads_or_all(n_threads);
        int64_t *lv = (int64_t *)calloc((size_t)nt * v->depth, sizeof(int64_t));
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
        for (int64_t r = 0; r < R; ++r) {
#ifdef _OPENMP
            int64_t *mylive = lv + (int64_t)omp_get_thread_num() * v->depth;
#else
            int64_t *mylive = lv;
#endif
            unit u = init[r];
            float t;
            int32_t id = kd_search(s, u.start_p, u.in_dir, &t, NULL);
            if (id < 0) {
                intersect_record(s, &u, accum, v->background, v->sample, NULL);
                continue;
            }
            trace_paths(s, v, &init[r], id, t, accum, mylive);
        }
        if (live) {
            live[0] = R;
            for (int k = 0; k < nt; ++k)
                for (int i = 1; i < v->depth; ++i) live[i] += lv[(int64_t)k * v->depth + i];
        }
        free(lv);
        free(init);
    }
    if (rgba)
        orc_get_image(accum, v->sample_x, v->sample_y, v->supersample_x, v->supersample_y,
                      v->sample, v->brightness, rgba);
    return 0;
}

/* Literal wavefront in seed_mode 0, exported so tests can check that the
 * depth-first engine above reorders the same arithmetic. */
int orc_render_wavefront(const orc_scene *s, const orc_view *v, int seed_mode, float *accum,
                         uint8_t *rgba, int64_t *live) {
    if (render_wavefront(s, v, seed_mode, accum, live)) return -1;
    if (rgba)
        orc_get_image(accum, v->sample_x, v->sample_y, v->supersample_x, v->supersample_y,
                      v->sample, v->brightness, rgba);
    return 0;
}
