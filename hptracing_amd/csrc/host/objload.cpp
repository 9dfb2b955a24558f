// OBJ / MTL parsing with the behaviour of the tinyobjloader 0.9.7 copy that
// hpTracing vendors (src/objloader/tiny_obj_loader.cpp), including its local
// changes (`Ke` emission :389-397, `Tf` as dissolve :425-429) and its quirks:
//   * `usemtl` drops the faces collected so far in the current group
//     (:629), so only faces after the last `usemtl` of a group survive;
//   * vertices are de-duplicated per shape on the (v, vn, vt) triple, in
//     first-use order of the triangle fan (:164-203, :242-265);
//   * normals / texcoords are appended only for corners that carry them,
//     so the streams can be shorter than the positions (:188-197).
// Deviations, chosen because the reference behaviour is undefined there:
// an unreadable mtllib and a vertex index out of range raise an error.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>

#include "host.h"

namespace hpt {
namespace {

struct Corner {
    int v, vt, vn;
};
// Cache key order: v, then vn, then vt (tiny_obj_loader.cpp:43-50).
struct CornerLess {
    bool operator()(const Corner &a, const Corner &b) const {
        if (a.v != b.v) return a.v < b.v;
        if (a.vn != b.vn) return a.vn < b.vn;
        return a.vt < b.vt;
    }
};

inline bool is_space(char c) { return c == ' ' || c == '\t'; }
inline bool is_eol(char c) { return c == '\r' || c == '\n' || c == '\0'; }

// 1-based to 0-based, 0 stays 0, negative = relative to the current count.
inline int to_index(int idx, int n) {
    if (idx > 0) return idx - 1;
    if (idx == 0) return 0;
    return n + idx;
}

// atof then narrow to float (two roundings, as the reference does).
inline float next_float(const char *&p) {
    p += std::strspn(p, " \t");
    float f = (float)std::atof(p);
    p += std::strcspn(p, " \t\r");
    return f;
}

inline int next_int(const char *&p) {
    p += std::strspn(p, " \t");
    int i = std::atoi(p);
    p += std::strcspn(p, " \t\r");
    return i;
}

Corner parse_corner(const char *&p, int nv, int nvn, int nvt) {
    Corner c{-1, -1, -1};
    c.v = to_index(std::atoi(p), nv);
    p += std::strcspn(p, "/ \t\r");
    if (p[0] != '/') return c;
    ++p;
    if (p[0] == '/') {  // v//vn
        ++p;
        c.vn = to_index(std::atoi(p), nvn);
        p += std::strcspn(p, "/ \t\r");
        return c;
    }
    c.vt = to_index(std::atoi(p), nvt);  // v/vt or v/vt/vn
    p += std::strcspn(p, "/ \t\r");
    if (p[0] != '/') return c;
    ++p;
    c.vn = to_index(std::atoi(p), nvn);
    p += std::strcspn(p, "/ \t\r");
    return c;
}

void reset_mtl(MtlEntry &m) {
    m.name.clear();
    m.ambient_texname.clear();
    m.diffuse_texname.clear();
    m.specular_texname.clear();
    m.normal_texname.clear();
    for (int i = 0; i < 3; ++i) {
        m.ambient[i] = m.diffuse[i] = m.specular[i] = 0.f;
        m.transmittance[i] = m.emission[i] = 0.f;
    }
    m.illum = 0;
    m.dissolve = 1.f;
    m.shininess = 1.f;
    m.ior = 1.f;
}

// Reads one line the way the reference loop does: at most 8191 characters,
// trailing '\n' / '\r' removed.  Returns false at end of input (also after
// an over-long line, which leaves the stream failed in the reference).
bool read_line(std::istream &in, std::string &line) {
    if (in.peek() == EOF) return false;
    std::getline(in, line);
    if (line.size() >= 8192) return false;
    if (!line.empty() && line.back() == '\n') line.pop_back();
    if (!line.empty() && line.back() == '\r') line.pop_back();
    return true;
}

inline bool key3(const char *t, char a, char b) {
    return t[0] == a && t[1] == b && is_space(t[2]);
}

void read_rgb(const char *p, float out[3]) {
    p += 2;
    out[0] = next_float(p);
    out[1] = next_float(p);
    out[2] = next_float(p);
}

// LoadMtl (tiny_obj_loader.cpp:278-476).
void load_mtl(const std::string &path, std::vector<MtlEntry> &mtls,
              std::map<std::string, int> &by_name) {
    std::ifstream in(path.c_str());
    if (!in) throw Error(HPT_ERR_IO, "cannot open material library [" + path + "]");
    by_name.clear();
    MtlEntry m;
    reset_mtl(m);
    std::string line;
    while (read_line(in, line)) {
        if (line.empty()) continue;
        const char *t = line.c_str();
        t += std::strspn(t, " \t");
        if (t[0] == '\0' || t[0] == '#') continue;
        if (std::strncmp(t, "newmtl", 6) == 0 && is_space(t[6])) {
            if (!m.name.empty()) {
                by_name.insert(std::make_pair(m.name, (int)mtls.size()));
                mtls.push_back(m);
            }
            reset_mtl(m);
            char name[4096] = {0};
            std::sscanf(t + 7, "%4095s", name);
            m.name = name;
            continue;
        }
        if (key3(t, 'K', 'a')) { read_rgb(t, m.ambient); continue; }
        if (key3(t, 'K', 'd')) { read_rgb(t, m.diffuse); continue; }
        if (key3(t, 'K', 's')) { read_rgb(t, m.specular); continue; }
        if (key3(t, 'K', 't')) { read_rgb(t, m.transmittance); continue; }
        if (key3(t, 'N', 'i')) { t += 2; m.ior = next_float(t); continue; }
        if (key3(t, 'K', 'e')) { read_rgb(t, m.emission); continue; }
        if (key3(t, 'N', 's')) { t += 2; m.shininess = next_float(t); continue; }
        if (std::strncmp(t, "illum", 5) == 0 && is_space(t[5])) {
            t += 6;
            m.illum = next_int(t);
            continue;
        }
        if (t[0] == 'd' && is_space(t[1])) { t += 1; m.dissolve = next_float(t); continue; }
        if (key3(t, 'T', 'r')) { t += 2; m.dissolve = next_float(t); continue; }
        if (key3(t, 'T', 'f')) { t += 2; m.dissolve = next_float(t); continue; }
        if (std::strncmp(t, "map_Ka", 6) == 0 && is_space(t[6])) { m.ambient_texname = t + 7; continue; }
        if (std::strncmp(t, "map_Kd", 6) == 0 && is_space(t[6])) { m.diffuse_texname = t + 7; continue; }
        if (std::strncmp(t, "map_Ks", 6) == 0 && is_space(t[6])) { m.specular_texname = t + 7; continue; }
        if (std::strncmp(t, "map_Ns", 6) == 0 && is_space(t[6])) { m.normal_texname = t + 7; continue; }
        // Unknown keys are ignored (the reference stores them in a map).
    }
    // The last material is always flushed (:471-473).
    by_name.insert(std::make_pair(m.name, (int)mtls.size()));
    mtls.push_back(m);
}

// exportFaceGroupToShape (tiny_obj_loader.cpp:225-276): fan-triangulate the
// group's faces into `shape` with a fresh per-shape vertex cache.
bool flush_group(ObjShape &shape, const std::vector<float> &v,
                 const std::vector<float> &vn, const std::vector<float> &vt,
                 const std::vector<std::vector<Corner>> &faces, int material,
                 const std::string &name) {
    if (faces.empty()) return false;
    std::map<Corner, uint32_t, CornerLess> cache;
    auto emit = [&](const Corner &c) -> uint32_t {
        auto it = cache.find(c);
        if (it != cache.end()) return it->second;
        if (c.v < 0 || (size_t)(3 * (int64_t)c.v + 2) >= v.size())
            throw Error(HPT_ERR_SCENE, "face references a missing vertex");
        shape.positions.push_back(v[3 * c.v + 0]);
        shape.positions.push_back(v[3 * c.v + 1]);
        shape.positions.push_back(v[3 * c.v + 2]);
        if (c.vn >= 0) {
            if ((size_t)(3 * (int64_t)c.vn + 2) >= vn.size())
                throw Error(HPT_ERR_SCENE, "face references a missing normal");
            shape.normals.push_back(vn[3 * c.vn + 0]);
            shape.normals.push_back(vn[3 * c.vn + 1]);
            shape.normals.push_back(vn[3 * c.vn + 2]);
        }
        if (c.vt >= 0) {
            if ((size_t)(2 * (int64_t)c.vt + 1) >= vt.size())
                throw Error(HPT_ERR_SCENE, "face references a missing texcoord");
            shape.texcoords.push_back(vt[2 * c.vt + 0]);
            shape.texcoords.push_back(vt[2 * c.vt + 1]);
        }
        uint32_t idx = (uint32_t)(shape.positions.size() / 3 - 1);
        cache[c] = idx;
        return idx;
    };
    for (const auto &face : faces) {
        if (face.size() < 2) continue;  // the reference reads face[1] regardless
        Corner c0 = face[0], c2 = face[1];
        for (size_t k = 2; k < face.size(); ++k) {
            Corner c1 = c2;
            c2 = face[k];
            uint32_t a = emit(c0), b = emit(c1), c = emit(c2);
            shape.indices.push_back(a);
            shape.indices.push_back(b);
            shape.indices.push_back(c);
            shape.material_ids.push_back(material);
        }
    }
    shape.name = name;
    return true;
}

}  // namespace

// LoadObj (tiny_obj_loader.cpp:495-723).
void load_obj(const std::string &path, const std::string &mtl_basepath,
              std::vector<ObjShape> &shapes, std::vector<MtlEntry> &mtls) {
    shapes.clear();
    std::ifstream in(path.c_str());
    if (!in) throw Error(HPT_ERR_IO, "Cannot open file [" + path + "]");

    std::vector<float> v, vn, vt;
    std::vector<std::vector<Corner>> faces;
    std::string name;
    std::map<std::string, int> mtl_by_name;
    int material = -1;
    ObjShape shape;

    std::string line;
    while (read_line(in, line)) {
        if (line.empty()) continue;
        const char *t = line.c_str();
        t += std::strspn(t, " \t");
        if (t[0] == '\0' || t[0] == '#') continue;

        if (t[0] == 'v' && is_space(t[1])) {
            t += 2;
            float x = next_float(t), y = next_float(t), z = next_float(t);
            v.push_back(x); v.push_back(y); v.push_back(z);
            continue;
        }
        if (key3(t, 'v', 'n')) {
            t += 3;
            float x = next_float(t), y = next_float(t), z = next_float(t);
            vn.push_back(x); vn.push_back(y); vn.push_back(z);
            continue;
        }
        if (key3(t, 'v', 't')) {
            t += 3;
            float x = next_float(t), y = next_float(t);
            vt.push_back(x); vt.push_back(y);
            continue;
        }
        if (t[0] == 'f' && is_space(t[1])) {
            t += 2;
            t += std::strspn(t, " \t");
            std::vector<Corner> face;
            while (!is_eol(t[0])) {
                face.push_back(parse_corner(t, (int)(v.size() / 3), (int)(vn.size() / 3),
                                            (int)(vt.size() / 2)));
                t += std::strspn(t, " \t\r");
            }
            faces.push_back(face);
            continue;
        }
        if (std::strncmp(t, "usemtl", 6) == 0 && is_space(t[6])) {
            char nm[4096] = {0};
            std::sscanf(t + 7, "%4095s", nm);
            faces.clear();
            auto it = mtl_by_name.find(nm);
            material = (it != mtl_by_name.end()) ? it->second : -1;
            continue;
        }
        if (std::strncmp(t, "mtllib", 6) == 0 && is_space(t[6])) {
            char nm[4096] = {0};
            std::sscanf(t + 7, "%4095s", nm);
            load_mtl(mtl_basepath + nm, mtls, mtl_by_name);
            continue;
        }
        if ((t[0] == 'g' || t[0] == 'o') && is_space(t[1])) {
            if (flush_group(shape, v, vn, vt, faces, material, name)) shapes.push_back(shape);
            shape = ObjShape();
            faces.clear();
            if (t[0] == 'g') {
                // names[1] of the whitespace-separated tag list, or "".
                const char *p = t + 1;
                p += std::strspn(p, " \t");
                size_t n = std::strcspn(p, " \t\r");
                name.assign(p, n);
            } else {
                char nm[4096] = {0};
                std::sscanf(t + 2, "%4095s", nm);
                name = nm;
            }
            continue;
        }
        // Unknown commands are ignored.
    }
    if (flush_group(shape, v, vn, vt, faces, material, name)) shapes.push_back(shape);
}

}  // namespace hpt
