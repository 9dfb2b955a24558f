/*
 * Float64 restatement of the path tracer -- TEST INFRASTRUCTURE.
 *
 * The same algorithm as oracle.c (kernel.cl's intersect / s1_run / s2_*
 * rules and TraceRunner::run's pass structure, with per-path LCG streams),
 * written a second time with every geometric and shading quantity in
 * double precision.  It renders the high-sample reference images against
 * which the float32 GPU renders are judged by RMSE; a bias in either
 * implementation shows up as a difference larger than the noise.
 * Scene arrays are the float32 inputs of oracle.h, widened on load.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

typedef struct { double x, y, z; } d3;

static inline d3 mk(double x, double y, double z) { d3 r = {x, y, z}; return r; }
static inline d3 ldf(orc_f3 p) { return mk(p.x, p.y, p.z); }
static inline d3 add(d3 a, d3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline d3 sub(d3 a, d3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline d3 mulv(d3 a, d3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline d3 scl(double s, d3 a) { return mk(s * a.x, s * a.y, s * a.z); }
static inline d3 neg(d3 a) { return mk(-a.x, -a.y, -a.z); }
static inline double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline d3 cross(d3 a, d3 b) {
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline double len(d3 a) { return sqrt(dot(a, a)); }
static inline d3 nrm(d3 a) { double l = len(a); return mk(a.x / l, a.y / l, a.z / l); }

#define PI_D 3.14159265358979323846

static int box_hit(d3 bs, d3 be, d3 sp, d3 d) {
    if (sp.x >= bs.x && sp.x <= be.x && sp.y >= bs.y && sp.y <= be.y && sp.z >= bs.z &&
        sp.z <= be.z)
        return 1;
    double lo[3] = {bs.x, bs.y, bs.z}, hi[3] = {be.x, be.y, be.z};
    double p[3] = {sp.x, sp.y, sp.z}, v[3] = {d.x, d.y, d.z};
    double mx = -INFINITY, mn = INFINITY;
    for (int k = 0; k < 3; ++k) {
        double t0 = (lo[k] - p[k]) / v[k], t1 = (hi[k] - p[k]) / v[k];
        double a = fmin(t0, t1), b = fmax(t0, t1);
        mx = fmax(mx, a);
        mn = fmin(mn, b);
    }
    if (mn <= 0) return 0;
    return mx <= mn;
}

static double tri_hit(d3 sp, d3 d, d3 pa, d3 pb, d3 pc) {
    d3 e1 = sub(pb, pa), e2 = sub(pc, pa);
    d3 P = cross(d, e2);
    double inv = 1.0 / dot(e1, P);
    d3 T = sub(sp, pa);
    double u = dot(T, P) * inv;
    if (u < 0 || u > 1) return -1;
    d3 Q = cross(T, e1);
    double v = dot(d, Q) * inv;
    if (v < 0 || u + v > 1) return -1;
    double t = dot(e2, Q) * inv;
    return t > 0 ? t : -1;
}

/* kdtree_intersect's search (pre-order walk, first 512 non-empty leaves). */
static int32_t cast(const orc_scene *s, d3 sp, d3 d, double *t_out) {
    int leaves = 0, node = 0, from_child = 0;
    int32_t best_id = -1;
    double best = -42;
    for (;;) {
        const orc_node *n = &s->nodes[node];
        int down = 0;
        if (!from_child && box_hit(ldf(n->box_start), ldf(n->box_end), sp, d)) {
            if (n->child < 0) {
                if (n->data >= 0 && leaves < ORC_MAX_MATCH_DATA) {
                    ++leaves;
                    const int32_t *list = s->leaf_data + n->data;
                    for (int k = 0; k < list[0]; ++k) {
                        const int32_t *g = &s->mesh[4 * (int64_t)list[k + 1]];
                        double r = tri_hit(sp, d, ldf(s->points[g[0]]), ldf(s->points[g[1]]),
                                           ldf(s->points[g[2]]));
                        if (r > 0 && (best < 0 || r < best)) {
                            best = r;
                            best_id = list[k + 1];
                        }
                    }
                }
            } else {
                down = 1;
            }
        }
        if (down) {
            from_child = 0;
            node = n->child;
        } else if (n->sibling >= 0) {
            from_child = 0;
            node = n->sibling;
        } else if (n->parent >= 0) {
            from_child = 1;
            node = n->parent;
        } else {
            break;
        }
    }
    *t_out = best;
    return best_id;
}

static inline uint32_t lcg(uint64_t *s) {
    *s = (*s * 0x5DEECE66DULL + 0xBULL) & ((1ULL << 48) - 1);
    return (uint32_t)(*s >> 16);
}
static inline double rnd(uint64_t *s) { return (double)lcg(s) / 4294967296.0; }
static inline d3 rnd3(uint64_t *s) {
    double a = rnd(s) * PI_D, b = rnd(s) * PI_D;
    return mk(cos(b) * cos(a), cos(b) * sin(a), sin(b));
}

static d3 texel(const orc_scene *s, int layer, int i, int j) {
    const uint8_t *p = s->textures + (((int64_t)layer * 512 + j) * 512 + i) * 4;
    return mk(p[0] / 255.0, p[1] / 255.0, p[2] / 255.0);
}

static d3 texture(const orc_scene *s, int layer, double su, double sv) {
    if (layer > s->n_textures - 1) layer = (int)s->n_textures - 1;
    double u = (su - floor(su)) * 512, v = (sv - floor(sv)) * 512;
    double fu = floor(u - 0.5), fv = floor(v - 0.5);
    if (!isfinite(fu) || !isfinite(fv)) return texel(s, layer, 0, 0);
    int i0 = (int)fu, j0 = (int)fv, i1 = i0 + 1, j1 = j0 + 1;
    if (i0 < 0) i0 += 512;
    if (j0 < 0) j0 += 512;
    if (i1 > 511) i1 -= 512;
    if (j1 > 511) j1 -= 512;
    double a = u - 0.5 - fu, b = v - 0.5 - fv;
    return add(add(scl((1 - a) * (1 - b), texel(s, layer, i0, j0)), scl(a * (1 - b), texel(s, layer, i1, j0))),
               add(scl((1 - a) * b, texel(s, layer, i0, j1)), scl(a * b, texel(s, layer, i1, j1))));
}

/* One path of one pass (s1_run then the chosen s2 rule per depth level). */
static void trace_path(const orc_scene *s, const orc_view *v, d3 o, d3 d, int32_t id, double t,
                       uint64_t seed, double *acc) {
    d3 strength = mk(1, 1, 1);
    const d3 bg = mk(v->background[0], v->background[1], v->background[2]);
    for (int depth = 0; depth < v->depth; ++depth) {
        if (depth > 0) {
            id = cast(s, o, d, &t);
            if (id < 0) {
                d3 val = mulv(bg, strength);
                acc[0] += val.x;
                acc[1] += val.y;
                acc[2] += val.z;
                return;
            }
        }
        const int32_t *g = &s->mesh[4 * (int64_t)id];
        d3 A = ldf(s->points[g[0]]), B = ldf(s->points[g[1]]), C = ldf(s->points[g[2]]);
        const orc_material *m = &s->materials[g[3]];
        d3 P = add(o, scl(t, d));
        double all = len(cross(sub(B, A), sub(C, A)));
        d3 f1 = sub(A, P), f2 = sub(B, P), f3 = sub(C, P);
        double a1 = len(cross(f2, f3)) / all, a2 = len(cross(f1, f3)) / all,
               a3 = len(cross(f1, f2)) / all;
        d3 n = s->n_normals == 0
                   ? nrm(cross(sub(B, A), sub(C, A)))
                   : add(add(scl(a1, ldf(s->normals[g[0]])), scl(a2, ldf(s->normals[g[1]]))),
                         scl(a3, ldf(s->normals[g[2]])));
        double uvx = 0, uvy = 0;
        if (m->ambient_texture_id >= 0 || m->diffuse_texture_id >= 0) {
            if (s->n_texcoords == 0) {
                double l = sqrt(P.x * P.x + P.y * P.y);
                uvx = P.x / l;
                uvy = P.y / l;
            } else {
                const float *ta = s->texcoords + 2 * (int64_t)g[0], *tb = s->texcoords + 2 * (int64_t)g[1],
                            *tc = s->texcoords + 2 * (int64_t)g[2];
                uvx = a1 * ta[0] + a2 * tb[0] + a3 * tc[0];
                uvy = a1 * ta[1] + a2 * tb[1] + a3 * tc[1];
            }
            uvy = 1 - uvy;
        }
        d3 amb = ldf(m->ambient);
        if (m->ambient_texture_id >= 0) amb = texture(s, m->ambient_texture_id, uvx, uvy);
        d3 res = scl(fabs(dot(n, d)), mulv(strength, add(amb, ldf(m->emission))));
        acc[0] += res.x;
        acc[1] += res.y;
        acc[2] += res.z;
        if (len(strength) < 1e-4 || depth == v->depth - 1) return;
        double r = rnd(&seed);
        int br;
        if (r < m->specular_possibility) {
            strength = mulv(strength, ldf(m->specular));
            br = 1;
        } else if (r < m->refract_possibility) {
            strength = scl(1 - m->dissolve, strength);
            br = 0;
        } else {
            d3 dif = ldf(m->diffuse);
            if (m->diffuse_texture_id >= 0) dif = texture(s, m->diffuse_texture_id, uvx, uvy);
            strength = mulv(strength, dif);
            br = (r < m->diffuse_possibility && v->disable_diffuse == 0) ? 2 : 3;
        }
        d3 nd;
        if (br == 0) {
            double ca = dot(d, neg(n));
            double rev = ca < 0 ? -1 : 1;
            double al = acos(ca);
            d3 p = scl(ca, n), q = nrm(add(d, p));
            double sb = sin(al) * pow(m->optical_density, -rev);
            nd = sb <= 1 ? add(scl(-rev * cos(asin(sb)), n), scl(sb, q)) : add(add(d, p), p);
        } else if (br == 1) {
            d3 rv = scl(1.0 / 2.0 / m->shininess, rnd3(&seed));
            nd = add(sub(d, scl(2 * dot(d, n), n)), rv);
        } else if (br == 2) {
            int front = dot(d, n) < 0;
            d3 p = rnd3(&seed);
            double dn = dot(p, n);
            if (dn < 0) {
                dn = -dn;
                if (front) p = neg(p);
            }
            strength = scl(dn, strength);
            nd = p;
        } else {
            if (s->n_lights == 0) return;
            int front = dot(d, n) < 0;
            const int32_t *l = &s->lights[4 * (int64_t)(lcg(&seed) % (uint32_t)s->n_lights)];
            d3 pa = ldf(s->points[l[0]]), pb = ldf(s->points[l[1]]), pc = ldf(s->points[l[2]]);
            double rx = rnd(&seed), ry = rnd(&seed);
            if (rx + ry > 1) {
                rx = 1 - rx;
                ry = 1 - ry;
            }
            d3 p = nrm(sub(add(add(pa, scl(rx, sub(pb, pa))), scl(ry, sub(pc, pa))), P));
            double dl = dot(p, n);
            if ((dl > 0) != front) return;
            strength = scl(fabs(dl), strength);
            nd = p;
        }
        o = add(P, scl(1e-5, nd));
        d = nd;
    }
}

/* Render with per-path streams orc_path_seed(seed, ray, pass).  accum:
 * W*ssx*H*ssy*3 sums over passes (like results_mem); accum_sq (may be NULL):
 * sums of the squared per-pass values, for standard errors. */
int orc64_render(const orc_scene *s, const orc_view *v, int n_threads, double *accum,
                 double *accum_sq) {
    if (v->height < 0) return -1;  /* pinhole camera only */
    const int sx = v->sample_x * v->supersample_x, sy = v->sample_y * v->supersample_y;
    const int64_t R = (int64_t)sx * sy;
    memset(accum, 0, sizeof(double) * 3 * R);
    if (accum_sq) memset(accum_sq, 0, sizeof(double) * 3 * R);
    const d3 up = mk(v->up[0], v->up[1], v->up[2]), right = mk(v->right[0], v->right[1], v->right[2]);
    const d3 org = mk(v->view[0], v->view[1], v->view[2]);
#ifdef _OPENMP
    int nt = n_threads > 0 ? n_threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 8) num_threads(nt)
#endif
    for (int64_t r = 0; r < R; ++r) {
        const int gx = (int)(r % sx), gy = (int)(r / sx);
        d3 nd = cross(up, right);
        nd = add(nd, scl((double)gy / sy * v->height - v->height * 0.5, up));
        nd = add(nd, scl((double)gx / sx * v->width - v->width * 0.5, right));
        const d3 d = nrm(nd);
        double t;
        int32_t id = cast(s, org, d, &t);
        double *acc = accum + 3 * r;
        double *acc2 = accum_sq ? accum_sq + 3 * r : NULL;
        if (id < 0) {
            for (int k = 0; k < 3; ++k) {
                const double b = (double)v->background[k];
                acc[k] = b * v->sample;
                if (acc2) acc2[k] = b * b * v->sample;
            }
            continue;
        }
        for (int p = 0; p < v->sample; ++p) {
            double pv[3] = {0, 0, 0};
            trace_path(s, v, org, d, id, t, orc_path_seed(v->seed, r, p), pv);
            for (int k = 0; k < 3; ++k) {
                acc[k] += pv[k];
                if (acc2) acc2[k] += pv[k] * pv[k];
            }
        }
    }
    return 0;
}
