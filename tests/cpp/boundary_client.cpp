// A program written against the C++ boundary (include/hptracing.hpp) the way
// the reference's CLI uses hp::KDTree and hp::TraceRunner
// (src/hptracing.cpp:246-249 builds them, :74-78 calls run()): it renders one
// view and writes the RGBA8 pixels raw to a file.
//
//   boundary_client OBJ MTL_BASE OUT W H ANGLE SUPERSAMPLE SPP DEPTH NO_DIFFUSE
//                   BRIGHTNESS SEED VX VY VZ UX UY UZ RX RY RZ BR BG BB
//
// Prints "nodes N leaf_ints M rays R" on success; exits 3 with the message
// of the hpt::Error it caught otherwise.
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include <hptracing.hpp>

int main(int argc, char **argv) {
    if (argc != 25) {
        std::fprintf(stderr, "usage: see the header of boundary_client.cpp\n");
        return 2;
    }
    const char *obj = argv[1], *mtl_base = argv[2], *out = argv[3];
    const int width = std::atoi(argv[4]), height = std::atoi(argv[5]);
    const float angle = std::strtof(argv[6], nullptr);
    const bool supersample = std::atoi(argv[7]) != 0;
    const int sample = std::atoi(argv[8]), depth = std::atoi(argv[9]);
    const bool no_diffuse = std::atoi(argv[10]) != 0;
    const float brightness = std::strtof(argv[11], nullptr);
    const unsigned long long seed = std::strtoull(argv[12], nullptr, 10);
    auto f3 = [&](int i) {
        return hpt::Float3{std::strtof(argv[i], nullptr), std::strtof(argv[i + 1], nullptr),
                           std::strtof(argv[i + 2], nullptr)};
    };
    const hpt::Float3 view_point = f3(13), up_dir = f3(16), right_dir = f3(19),
                      background_color = f3(22);
    try {
        auto scene = std::make_unique<hpt::KDTree>(obj, mtl_base);
        auto data = scene->getData();
        auto runner = std::make_unique<hpt::TraceRunner>(scene);
        runner->set_seed(seed);
        std::vector<unsigned char> pixels((size_t)width * height * 4);
        runner->run(pixels.data(), view_point, up_dir, right_dir,
                    float(width) / float(height) * angle, angle, width, height,
                    supersample ? 2 : 1, supersample ? 2 : 1, background_color, sample, depth,
                    no_diffuse ? 1 : 0, brightness);
        FILE *f = std::fopen(out, "wb");
        if (!f || std::fwrite(pixels.data(), 1, pixels.size(), f) != pixels.size()) {
            std::fprintf(stderr, "cannot write %s\n", out);
            return 4;
        }
        std::fclose(f);
        std::printf("nodes %zu leaf_ints %zu rays %lld\n", data.first.size(), data.second.size(),
                    (long long)runner->stats().rays_cast);
    } catch (const hpt::Error &e) {
        std::fprintf(stderr, "hpt::Error %d: %s\n", e.code, e.what());
        return 3;
    }
    return 0;
}
