// Scene assembly with the behaviour of hp::Scene (src/hp/scene/base.cpp):
// materials and their branch probabilities, texture layers, flat
// point/normal/texcoord/mesh arrays and the 256-entry emitter table.
//
// The reference evaluates vector norms and dot products with Eigen 3
// (Vec = Eigen::Matrix<float,3,1>, common.h:39-40).  Eigen reduces a
// fixed-size 3-vector with its unrolled pairwise reducer, which evaluates
// a0 + (a1 + a2); `sum3` below keeps that order so the float results match.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>

#include "host.h"

namespace hpt {
namespace {

inline float sum3(float a0, float a1, float a2) { return a0 + (a1 + a2); }
inline float dot3(const float *a, const float *b) {
    return sum3(a[0] * b[0], a[1] * b[1], a[2] * b[2]);
}
inline float norm3(const float *a) { return std::sqrt(dot3(a, a)); }

}  // namespace

// Scene::Scene material conversion (scene/base.cpp:47-96).
hpt_material material_from_mtl(const float ka[3], const float kd[3], const float ks[3],
                               const float ke[3], float ior, float dissolve,
                               float shininess) {
    hpt_material x;
    std::memset(&x, 0, sizeof(x));
    for (int i = 0; i < 3; ++i) {
        x.emission[i] = ke[i];
        x.ambient[i] = ka[i];
        x.diffuse[i] = kd[i];
        x.specular[i] = ks[i];
    }
    x.optical_density = ior;
    x.dissolve = dissolve;
    x.shininess = shininess;
    x.diffuse_texture_id = -1;
    x.ambient_texture_id = -1;
    // DIFFUSE_FACTOR 0.9f; the 1e-4 is a double literal, so the sum is
    // formed in float, widened, offset in double and narrowed again.
    const float specular_length = norm3(ks);
    const float diffuse_length = norm3(kd);
    const float refract_length = 1.0f - dissolve;
    const float sum = (float)((double)(specular_length + diffuse_length + refract_length) + 1e-4);
    x.specular_possibility = specular_length / sum;
    x.refract_possibility = refract_length / sum + x.specular_possibility;
    x.diffuse_possibility = diffuse_length * 0.9f / sum + x.refract_possibility;
    return x;
}

// ---------------------------------------------------------------- images
namespace {

// Reads the next PNM header integer, skipping whitespace and '#' comments.
bool pnm_int(std::istream &in, int &out) {
    int c = in.get();
    while (c != EOF) {
        if (c == '#') {
            while (c != EOF && c != '\n') c = in.get();
        } else if (!std::isspace(c)) {
            break;
        }
        c = in.get();
    }
    if (c == EOF || !std::isdigit(c)) return false;
    long v = 0;
    while (c != EOF && std::isdigit(c)) {
        v = v * 10 + (c - '0');
        if (v > (1L << 30)) return false;
        c = in.get();
    }
    out = (int)v;
    return true;
}

}  // namespace

Image load_pnm(const std::string &path) {
    std::ifstream in(path.c_str(), std::ios::binary);
    if (!in) throw Error(HPT_ERR_IO, "cannot open texture [" + path + "]");
    char m0 = 0, m1 = 0;
    in.get(m0);
    in.get(m1);
    if (m0 != 'P' || (m1 != '2' && m1 != '3' && m1 != '5' && m1 != '6'))
        throw Error(HPT_ERR_IO, "texture [" + path + "]: not a PNM file");
    int w = 0, h = 0, maxval = 0;
    if (!pnm_int(in, w) || !pnm_int(in, h) || !pnm_int(in, maxval) || w <= 0 || h <= 0 ||
        maxval <= 0 || maxval > 65535)
        throw Error(HPT_ERR_IO, "texture [" + path + "]: bad PNM header");
    Image img;
    img.w = w;
    img.h = h;
    img.c = (m1 == '3' || m1 == '6') ? 3 : 1;
    img.data.assign((size_t)w * h * img.c, 0.f);
    const bool ascii = (m1 == '2' || m1 == '3');
    const bool wide = maxval > 255;
    for (int y = 0; y < h; ++y)
        for (int x = 0; x < w; ++x)
            for (int ch = 0; ch < img.c; ++ch) {
                int val = 0;
                if (ascii) {
                    if (!pnm_int(in, val)) throw Error(HPT_ERR_IO, "texture [" + path + "]: truncated");
                } else if (wide) {
                    int hi = in.get(), lo = in.get();
                    if (lo == EOF) throw Error(HPT_ERR_IO, "texture [" + path + "]: truncated");
                    val = (hi << 8) | lo;
                } else {
                    int b = in.get();
                    if (b == EOF) throw Error(HPT_ERR_IO, "texture [" + path + "]: truncated");
                    val = b;
                }
                img.at(x, y, ch) = (float)val;
            }
    return img;
}

// CImg<float>(path) as the reference builds it (no libpng or libjpeg): PNM
// files by extension through CImg's own reader; every other file through
// ImageMagick, which recognises the content (CImg.h:41305 load_other ->
// :41063 load_imagemagick_external).  PNG and PNM content are decoded
// here; anything else (JPEG and the rest of ImageMagick's formats) is an
// error.
Image load_texture(const std::string &path) {
    std::string ext;
    const size_t dot = path.rfind('.');
    if (dot != std::string::npos) ext = path.substr(dot + 1);
    for (auto &ch : ext) ch = (char)std::tolower((unsigned char)ch);
    if (ext == "ppm" || ext == "pgm" || ext == "pnm") return load_pnm(path);
    unsigned char head[8] = {};
    {
        std::ifstream in(path.c_str(), std::ios::binary);
        if (!in) throw Error(HPT_ERR_IO, "cannot open texture [" + path + "]");
        in.read(reinterpret_cast<char *>(head), 8);
    }
    if (head[0] == 0x89 && head[1] == 'P' && head[2] == 'N' && head[3] == 'G') return load_png(path);
    if (head[0] == 'P' && (head[1] == '2' || head[1] == '3' || head[1] == '5' || head[1] == '6'))
        return load_pnm(path);
    if (head[0] == 0xFF && head[1] == 0xD8 && head[2] == 0xFF) return load_jpeg(path);
    throw Error(HPT_ERR_IO, "texture [" + path + "]: only PNG, JPEG and PNM images are decoded");
}

// CImg<float>::normalize(0, 255) (CImg.h:19093-19105), then
// resize(512, 512, 1, 3) with nearest neighbour (interpolation 1,
// CImg.h:21063-21138: source index floor(dst * src_size / dst_size) on every
// axis, channels included), then the float -> uint8_t conversion of
// vector<uint8_t>::push_back and an alpha of 0 (scene/base.cpp:101-108).
void append_texture_layer(const Image &src, std::vector<uint8_t> &out) {
    Image img = src;
    float m = img.data[0], M = img.data[0];
    for (float v : img.data) {
        m = std::min(m, v);
        M = std::max(M, v);
    }
    if (m == M) {
        std::fill(img.data.begin(), img.data.end(), 0.f);
    } else if (m != 0.f || M != 255.f) {
        const float fm = m, fM = M;
        for (float &v : img.data) v = (v - fm) / (fM - fm) * 255.f + 0.f;
    }
    const int S = HPT_TEXTURE_SIZE;
    const size_t base = out.size();
    out.resize(base + (size_t)S * S * 4);
    for (int y = 0; y < S; ++y) {
        const int sy = (int)((uint64_t)y * img.h / S);
        for (int x = 0; x < S; ++x) {
            const int sx = (int)((uint64_t)x * img.w / S);
            uint8_t *px = &out[base + ((size_t)y * S + x) * 4];
            for (int k = 0; k < 3; ++k) {
                const int sc = (int)((uint64_t)k * img.c / 3);
                px[k] = (uint8_t)img.at(sx, sy, sc);
            }
            px[3] = 0;
        }
    }
}

// ----------------------------------------------------------------- scene
void scene_from_obj(const std::string &path, const std::string &mtl_basepath,
                    HostScene &scene) {
    std::vector<ObjShape> shapes;
    std::vector<MtlEntry> mtls;
    load_obj(path, mtl_basepath, shapes, mtls);

    for (const auto &m : mtls) {
        hpt_material x = material_from_mtl(m.ambient, m.diffuse, m.specular, m.emission, m.ior,
                                           m.dissolve, m.shininess);
        auto texture_id = [&](const std::string &name) -> int32_t {
            auto found = std::find(scene.texture_names.begin(), scene.texture_names.end(), name);
            if (found != scene.texture_names.end()) return (int32_t)(found - scene.texture_names.begin());
            scene.texture_names.push_back(name);
            return (int32_t)scene.texture_names.size() - 1;
        };
        if (!m.diffuse_texname.empty()) x.diffuse_texture_id = texture_id(m.diffuse_texname);
        if (!m.ambient_texname.empty()) x.ambient_texture_id = texture_id(m.ambient_texname);
        scene.materials.push_back(x);
    }
    for (const auto &name : scene.texture_names)
        append_texture_layer(load_texture(mtl_basepath + name), scene.texture_data);

    for (const auto &shape : shapes) {
        const int64_t base = scene.n_points();
        for (size_t i = 0; i + 2 < shape.positions.size(); i += 3) {
            scene.points.insert(scene.points.end(),
                                {shape.positions[i], shape.positions[i + 1], shape.positions[i + 2], 0.f});
        }
        for (size_t i = 0; i + 2 < shape.normals.size(); i += 3) {
            scene.normals.insert(scene.normals.end(),
                                 {shape.normals[i], shape.normals[i + 1], shape.normals[i + 2], 0.f});
        }
        for (size_t i = 0; i + 1 < shape.texcoords.size(); i += 2) {
            scene.texcoords.push_back(shape.texcoords[i]);
            scene.texcoords.push_back(shape.texcoords[i + 1]);
        }
        for (size_t i = 0; i + 2 < shape.indices.size(); i += 3) {
            const int mat = shape.material_ids[i / 3];
            if (mat < 0 || mat >= (int)scene.materials.size())
                throw Error(HPT_ERR_SCENE, "shape '" + shape.name + "' has faces without a known material (usemtl)");
            scene.mesh.insert(scene.mesh.end(), {(int32_t)(shape.indices[i] + base),
                                                 (int32_t)(shape.indices[i + 1] + base),
                                                 (int32_t)(shape.indices[i + 2] + base), mat});
        }
    }
}

// registerGeometry + finishRegister (scene/base.cpp:160-198).
void build_light_table(HostScene &scene) {
    scene.lights.clear();
    float total = 0.f;
    std::map<float, int> by_total;
    const int64_t nt = scene.n_triangles();
    for (int64_t g = 0; g < nt; ++g) {
        const int32_t *tri = &scene.mesh[4 * g];
        const hpt_material &mat = scene.materials[tri[3]];
        const float emission_norm = norm3(mat.emission);
        if (!(emission_norm > 0)) continue;
        const float *pa = &scene.points[4 * (int64_t)tri[0]];
        const float *pb = &scene.points[4 * (int64_t)tri[1]];
        const float *pc = &scene.points[4 * (int64_t)tri[2]];
        float dx[3], dy[3];
        for (int i = 0; i < 3; ++i) {
            dx[i] = pb[i] - pa[i];
            dy[i] = pc[i] - pa[i];
        }
        const float dot = dot3(dx, dy);
        const float ndx = norm3(dx), ndy = norm3(dy);
        float ux[3], uy[3];
        for (int i = 0; i < 3; ++i) {
            ux[i] = dx[i] / ndx;
            uy[i] = dy[i] / ndy;
        }
        const float cosine = dot3(ux, uy);
        const float area = dot / cosine * std::sqrt(1 - cosine * cosine);
        const float val = area * emission_norm;
        total += val;
        by_total[total] = (int)g;
    }
    if (by_total.empty()) return;  // "No emission material in scene": empty table
    for (int i = 0; i < HPT_LIGHT_TABLE_SIZE; ++i) {
        const float val = (float)i / (float)HPT_LIGHT_TABLE_SIZE * total;
        auto target = by_total.upper_bound(val);
        if (target == by_total.end()) target = by_total.begin();
        const int32_t *tri = &scene.mesh[4 * (int64_t)target->second];
        scene.lights.insert(scene.lights.end(), {tri[0], tri[1], tri[2], tri[3]});
    }
}

}  // namespace hpt
