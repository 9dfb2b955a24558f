// Internal declarations of the CPU side: OBJ/MTL parsing, scene assembly,
// textures, emitter table and the KD-tree builder.  These keep the
// behaviour of the reference's CPU code (src/objloader, src/hp/scene) so
// that the GPU path receives byte-identical inputs.
#pragma once

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "hptracing.h"

namespace hpt {

// Error carrying an hpt status code; caught at the C boundary.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

// ------------------------------------------------------------ OBJ / MTL
// Parsed MTL entry (fields of tinyobj::material_t, tiny_obj_loader.h:15-35).
struct MtlEntry {
    std::string name;
    float ambient[3], diffuse[3], specular[3], transmittance[3], emission[3];
    float shininess, ior, dissolve;
    int illum;
    std::string ambient_texname, diffuse_texname, specular_texname, normal_texname;
};

// One shape: per-shape de-duplicated vertex streams and triangle indices
// (tinyobj::shape_t / mesh_t, tiny_obj_loader.h:37-50).
struct ObjShape {
    std::string name;
    std::vector<float> positions, normals, texcoords;
    std::vector<uint32_t> indices;
    std::vector<int> material_ids;
};

// Parse an OBJ file and its mtllib.  Throws hpt::Error.
void load_obj(const std::string &path, const std::string &mtl_basepath,
              std::vector<ObjShape> &shapes, std::vector<MtlEntry> &mtls);

// ---------------------------------------------------------------- images
// Decode a binary or ASCII PNM (P2/P3/P5/P6) file into planar float
// channels, as CImg's native loader does.  decode_png / load_png give the
// channels CImg receives for a PNG from ImageMagick's PNM output (gray: 1,
// colour and palette: 3, alpha dropped, samples as stored); decode_jpeg /
// load_jpeg give the samples of libjpeg's default decode (gray: 1, colour:
// 3) that ImageMagick prints for a JPEG; load_texture picks the reader as
// CImg<float>(path) does.  All throw hpt::Error.
struct Image {
    int w = 0, h = 0, c = 0;
    std::vector<float> data;  // planar: data[(ch * h + y) * w + x]
    float &at(int x, int y, int ch) { return data[((size_t)ch * h + y) * w + x]; }
    float at(int x, int y, int ch) const { return data[((size_t)ch * h + y) * w + x]; }
};
Image load_pnm(const std::string &path);
Image decode_png(const std::vector<uint8_t> &file, const std::string &name);
Image load_png(const std::string &path);
Image decode_jpeg(const std::vector<uint8_t> &file, const std::string &name);
Image load_jpeg(const std::string &path);
Image load_texture(const std::string &path);
// CImg::normalize(0,255) then resize(512,512,1,3) with nearest neighbour,
// then truncation to uint8 with alpha 0 (scene/base.cpp:98-110).
void append_texture_layer(const Image &img, std::vector<uint8_t> &out);

// ----------------------------------------------------------------- scene
struct HostScene {
    std::vector<hpt_material> materials;
    std::vector<float> points;     // 4 floats per point (cl_float3 layout)
    std::vector<float> normals;    // 4 floats per normal
    std::vector<float> texcoords;  // 2 floats per texcoord
    std::vector<int32_t> mesh;     // int4 per triangle
    std::vector<int32_t> lights;   // int4 per entry (0 or 256 entries)
    std::vector<std::string> texture_names;
    std::vector<uint8_t> texture_data;

    std::vector<hpt_kd_node> nodes;
    std::vector<int32_t> leaf_data;
    double build_seconds = 0.0;

    int64_t n_points() const { return (int64_t)points.size() / 4; }
    int64_t n_triangles() const { return (int64_t)mesh.size() / 4; }
};

hpt_material material_from_mtl(const float ka[3], const float kd[3],
                               const float ks[3], const float ke[3], float ior,
                               float dissolve, float shininess);

// Scene::Scene (scene/base.cpp:37-158) without the KD build.
void scene_from_obj(const std::string &path, const std::string &mtl_basepath,
                    HostScene &scene);
// registerGeometry + finishRegister (scene/base.cpp:160-198).
void build_light_table(HostScene &scene);
// KDTree ctor + getData (kdtree.cpp:235-292).
void build_kdtree(HostScene &scene, int n_threads);
// The size of the tree build_kdtree would make, without keeping it.
struct KdCounts {
    int64_t nodes = 0, leaves = 0, entries = 0, max_leaf = 0;
    int max_depth = 0;
    double seconds = 0;
};
KdCounts count_kdtree(const std::vector<float> &points, const std::vector<int32_t> &mesh,
                      int n_threads);

}  // namespace hpt
