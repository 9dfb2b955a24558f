// hptracing command line over the C ABI, with the options, defaults and
// argument mapping of the reference CLI (src/hptracing.cpp:213-298):
//   -i/--input, --mtl, --width 500, --height 500, --view 0,0,-500,
//   --up 0,1,0, --right 1,0,0, --angle 1.0 (-1 = equirectangular),
//   -x/--supersample (2x2), -s/--sample 10, -d/--depth 6, --brightness 1,
//   --no-diffuse, --background 0,0,0, -o/--output.
// run() receives W/H*angle and angle as the angular extents
// (hptracing.cpp:74-78); the saved image is mirrored in y (:289-295).
// Output formats: .ppm (P6) and .png (stored deflate), written here because
// the reference's CImg hands PNG to an external converter.
// Added: --seed (the reference seeds from the clock).  Without -o the
// reference opens a GLUT window; this build has no window and exits with an
// error instead.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "hptracing.h"

namespace {

// str2float3 (hptracing.cpp:31-39): atof, then continue after the next
// comma; without a comma the same value repeats.
void str2float3(std::string s, float out[3]) {
    for (int i = 0; i < 3; ++i) {
        out[i] = (float)std::atof(s.c_str());
        if (i < 2) s = s.substr(s.find(',') + 1);
    }
}

uint32_t crc32(const uint8_t *p, size_t n, uint32_t c = 0xFFFFFFFFu) {
    static uint32_t table[256];
    static bool init = false;
    if (!init) {
        for (uint32_t i = 0; i < 256; ++i) {
            uint32_t v = i;
            for (int k = 0; k < 8; ++k) v = (v & 1) ? 0xEDB88320u ^ (v >> 1) : v >> 1;
            table[i] = v;
        }
        init = true;
    }
    for (size_t i = 0; i < n; ++i) c = table[(c ^ p[i]) & 0xFF] ^ (c >> 8);
    return c;
}

void put32(std::vector<uint8_t> &o, uint32_t v) {
    o.push_back(v >> 24);
    o.push_back(v >> 16);
    o.push_back(v >> 8);
    o.push_back(v);
}

void chunk(std::vector<uint8_t> &png, const char *type, const std::vector<uint8_t> &data) {
    put32(png, (uint32_t)data.size());
    std::vector<uint8_t> body(type, type + 4);
    body.insert(body.end(), data.begin(), data.end());
    png.insert(png.end(), body.begin(), body.end());
    put32(png, crc32(body.data(), body.size()) ^ 0xFFFFFFFFu);
}

bool write_png(const std::string &path, const std::vector<uint8_t> &rgb, int w, int h) {
    std::vector<uint8_t> raw;
    raw.reserve((size_t)h * (3 * w + 1));
    for (int y = 0; y < h; ++y) {
        raw.push_back(0);
        raw.insert(raw.end(), rgb.begin() + (size_t)y * 3 * w, rgb.begin() + (size_t)(y + 1) * 3 * w);
    }
    std::vector<uint8_t> z = {0x78, 0x01};
    uint32_t a = 1, b = 0;
    for (uint8_t c : raw) {
        a = (a + c) % 65521;
        b = (b + a) % 65521;
    }
    for (size_t off = 0; off < raw.size() || off == 0; off += 65535) {
        const size_t n = std::min<size_t>(65535, raw.size() - off);
        z.push_back(off + n >= raw.size() ? 1 : 0);
        z.push_back(n & 0xFF);
        z.push_back(n >> 8);
        z.push_back(~n & 0xFF);
        z.push_back((~n >> 8) & 0xFF);
        z.insert(z.end(), raw.begin() + off, raw.begin() + off + n);
        if (n == 0) break;
    }
    put32(z, (b << 16) | a);
    std::vector<uint8_t> png = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
    std::vector<uint8_t> ihdr;
    put32(ihdr, w);
    put32(ihdr, h);
    ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});
    chunk(png, "IHDR", ihdr);
    chunk(png, "IDAT", z);
    chunk(png, "IEND", {});
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(png.data(), 1, png.size(), f) == png.size();
    return std::fclose(f) == 0 && ok;
}

bool write_ppm(const std::string &path, const std::vector<uint8_t> &rgb, int w, int h) {
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    std::fprintf(f, "P6\n%d %d\n255\n", w, h);
    const bool ok = std::fwrite(rgb.data(), 1, rgb.size(), f) == rgb.size();
    return std::fclose(f) == 0 && ok;
}

int usage(const char *argv0) {
    std::fprintf(stderr,
                 "usage: %s -i scene.obj -o out.png [--mtl dir/] [--width 500] [--height 500]\n"
                 "  [--view 0,0,-500] [--up 0,1,0] [--right 1,0,0] [--angle 1.0 | -1 for VR]\n"
                 "  [-x] [-s 10] [-d 6] [--brightness 1] [--no-diffuse] [--background 0,0,0]\n"
                 "  [--seed N] [--progressive PREFIX] [--tune]\n"
                 "  --tune runs hpt_autotune on the view before rendering (launch forms\n"
                 "  only; the image is the same)\n"
                 "  --progressive writes PREFIX<passes>.png each time the render server\n"
                 "  publishes more sample passes (the refining frames the original tool's\n"
                 "  window shows); -o is then optional\n", argv0);
    return 2;
}

}  // namespace

int main(int argc, char **argv) {
    std::string input, mtl, output, progressive, view = "0,0,-500", up = "0,1,0",
                                                 right = "1,0,0", background = "0,0,0";
    int width = 500, height = 500, sample = 10, depth = 6;
    float angle = 1.0f, brightness = 1.0f;
    bool supersample = false, no_diffuse = false, tune = false;
    unsigned long long seed = 1;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto next = [&](std::string &dst) {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "%s needs a value\n", a.c_str());
                std::exit(2);
            }
            dst = argv[++i];
        };
        std::string v;
        if (a == "-i" || a == "--input") next(input);
        else if (a == "--mtl") next(mtl);
        else if (a == "-o" || a == "--output") next(output);
        else if (a == "--progressive") next(progressive);
        else if (a == "--view") next(view);
        else if (a == "--up") next(up);
        else if (a == "--right") next(right);
        else if (a == "--background") next(background);
        else if (a == "--width") { next(v); width = std::atoi(v.c_str()); }
        else if (a == "--height") { next(v); height = std::atoi(v.c_str()); }
        else if (a == "-s" || a == "--sample") { next(v); sample = std::atoi(v.c_str()); }
        else if (a == "-d" || a == "--depth") { next(v); depth = std::atoi(v.c_str()); }
        else if (a == "--angle") { next(v); angle = (float)std::atof(v.c_str()); }
        else if (a == "--brightness") { next(v); brightness = (float)std::atof(v.c_str()); }
        else if (a == "--seed") { next(v); seed = std::strtoull(v.c_str(), nullptr, 10); }
        else if (a == "-x" || a == "--supersample") supersample = true;
        else if (a == "--no-diffuse") no_diffuse = true;
        else if (a == "--tune") tune = true;
        else if (a == "-h" || a == "--help") return usage(argv[0]);
        else {
            std::fprintf(stderr, "unknown option %s\n", a.c_str());
            return usage(argv[0]);
        }
    }
    if (input.empty()) return usage(argv[0]);
    if (output.empty() && progressive.empty()) {
        std::fprintf(stderr, "no -o given: the interactive window of the original tool is not "
                             "part of this build (--progressive writes its refining frames)\n");
        return 2;
    }
    if (mtl.empty()) {  // hptracing.cpp:238-244
        const size_t pos = input.rfind('/');
        if (pos != std::string::npos) mtl = input.substr(0, pos + 1);
    }

    hpt_host_scene *host = nullptr;
    if (hpt_host_load_obj(input.c_str(), mtl.c_str(), 0, &host) != HPT_OK) {
        std::fprintf(stderr, "loading %s: %s\n", input.c_str(), hpt_host_last_error());
        return 1;
    }
    std::fprintf(stderr, "Build KDTree done (%.2fs)\n", hpt_host_build_seconds(host));
    hpt_scene_desc desc;
    hpt_host_describe(host, &desc);
    hpt_scene *scene = nullptr;
    if (hpt_scene_create(&desc, &scene) != HPT_OK) {
        std::fprintf(stderr, "device scene: %s\n", hpt_last_error());
        return 1;
    }
    hpt_view v;
    std::memset(&v, 0, sizeof(v));
    str2float3(view, v.view);
    str2float3(up, v.up);
    str2float3(right, v.right);
    str2float3(background, v.background);
    v.width = (float)width / (float)height * angle;
    v.height = angle;
    v.sample_x = width;
    v.sample_y = height;
    v.supersample_x = v.supersample_y = supersample ? 2 : 1;
    v.sample = sample;
    v.depth = depth;
    v.disable_diffuse = no_diffuse ? 1 : 0;
    v.brightness = brightness;
    v.seed = seed;
    if (tune) {
        hpt_tuning chosen;
        if (hpt_autotune(scene, &v, 0, &chosen) != HPT_OK) {
            std::fprintf(stderr, "autotune: %s\n", hpt_last_error());
            return 1;
        }
        std::fprintf(stderr, "Tuned: fetch %d, %d cast blocks per CU, refill %d, node steps %d\n",
                     chosen.fetch, chosen.cast_blocks_per_cu, chosen.refill_threshold,
                     chosen.node_steps);
    }
    std::vector<uint8_t> rgba((size_t)width * height * 4);
    // hptracing.cpp:285-295: the image is written flipped in y.
    auto save = [&](const std::string &path) {
        std::vector<uint8_t> rgb((size_t)width * height * 3);
        for (int y = 0; y < height; ++y)
            for (int x = 0; x < width; ++x)
                for (int k = 0; k < 3; ++k)
                    rgb[((size_t)(height - 1 - y) * width + x) * 3 + k] =
                        rgba[((size_t)y * width + x) * 4 + k];
        const bool png = path.size() > 4 && path.substr(path.size() - 4) == ".png";
        const bool ok = png ? write_png(path, rgb, width, height) : write_ppm(path, rgb, width, height);
        if (!ok) std::fprintf(stderr, "cannot write %s\n", path.c_str());
        return ok;
    };
    bool ok = true;
    if (progressive.empty()) {
        hpt_render_stats st;
        if (hpt_render(scene, &v, rgba.data(), nullptr, &st) != HPT_OK) {
            std::fprintf(stderr, "render: %s\n", hpt_last_error());
            return 1;
        }
        std::fprintf(stderr, "Render done: %.3fs, %lld rays, %.1f Mrays/s\n", st.seconds,
                     (long long)st.rays_cast, st.rays_cast / st.seconds / 1e6);
        ok = save(output);
    } else {
        // The view as one progressive request of a render server whose
        // arena holds it (about 1.3 KB per camera ray).
        const int64_t rays = (int64_t)width * height * (supersample ? 4 : 1);
        hpt_server_opts opts{rays * 2048 + ((int64_t)64 << 20)};
        hpt_server *srv = nullptr;
        int64_t id = 0;
        if (hpt_server_create(scene, &opts, &srv) != HPT_OK ||
            hpt_server_submit_ex(srv, &v, nullptr, nullptr, HPT_SERVE_PROGRESSIVE, &id) != HPT_OK) {
            std::fprintf(stderr, "server: %s\n", hpt_last_error());
            return 1;
        }
        int32_t passes = 0;
        while (ok && passes < sample) {
            int32_t got = 0;
            if (hpt_server_peek(srv, id, passes + 1, -1.0, rgba.data(), &got) != HPT_OK) {
                std::fprintf(stderr, "server: %s\n", hpt_last_error());
                return 1;
            }
            if (got <= passes) break;  // the request ended
            passes = got;
            ok = save(progressive + std::to_string(passes) + ".png");
            std::fprintf(stderr, "%d of %d samples\n", passes, sample);
        }
        hpt_request_info info;
        const int rc = hpt_server_wait(srv, id, -1.0, &info);
        hpt_server_destroy(srv);
        if (rc != HPT_OK) {
            std::fprintf(stderr, "server: %s\n", hpt_last_error());
            return 1;
        }
        std::fprintf(stderr, "Render done: %.3fs\n", info.latency_ms / 1e3);
        if (ok && !output.empty()) ok = save(output);
    }
    hpt_scene_destroy(scene);
    hpt_host_free(host);
    return ok ? 0 : 1;
}
