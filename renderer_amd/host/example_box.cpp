/*
 * example_box.cpp — renders the box scene (5 quads, area light under the
 * ceiling, diffuse and mirror spheres) through the C++ host interface:
 * SceneBuilder -> GpuUniDiPathTracingIntegrator -> PerPixelRenderer.
 *
 *   example_box <out.raw> [width height spp_x spp_y mode max_depth seed]
 *   example_box <out.raw> --scene <file.scene> [spp_x spp_y mode max_depth seed]
 *   example_box <out.raw> --mlt <metropolis|metropolisflt|energyredist> [width height mutations seed]
 *   example_box <out.raw> --rays <rays.raw> [mode max_depth key seed]
 *
 * The second form reads the scene and its camera from a scene file
 * (SceneFile, the reference's scene language); the third renders the box
 * with MetropolisRenderer or EnergyRedistributionRenderer.  Writes
 * height * width * 3 float32 RGB and prints the mean radiance.  The fourth
 * integrates the rays of rays.raw (8 float32 per ray: origin, range start,
 * direction, range end) in the box one by one through the integrator's
 * per-ray entry and writes 3 float32 per ray.
 */
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "gpu_integrator.h"

static int write_film(const char* path, const std::vector<float>& film)
{
    double sum = 0;
    for (float v : film) sum += v;
    FILE* f = std::fopen(path, "wb");
    if (!f || std::fwrite(film.data(), sizeof(float), film.size(), f) != film.size()) {
        std::fprintf(stderr, "cannot write %s\n", path);
        if (f) std::fclose(f);
        return 1;
    }
    std::fclose(f);
    std::printf("mean %.6f\n", sum / film.size());
    return 0;
}

/* the box of renderer_amd/scenes.py box_scene */
static rtgpu::Camera box(rtgpu::SceneBuilder& s, int w, int h)
{
    int grey = s.diffuse({0.75f, 0.75f, 0.75f});
    int red = s.diffuse({0.63f, 0.06f, 0.05f});
    int green = s.diffuse({0.14f, 0.45f, 0.09f});
    int mirror = s.mirror();
    s.areaLight({0.375f, 0.999f, 0.375f}, {0.25f, 0, 0}, {0, 0, 0.25f}, {17, 12, 4});
    s.quad({0, 0, 0}, {1, 0, 0}, {1, 0, 1}, {0, 0, 1}, grey);
    s.quad({0, 1, 0}, {0, 1, 1}, {1, 1, 1}, {1, 1, 0}, grey);
    s.quad({0, 0, 0}, {0, 1, 0}, {1, 1, 0}, {1, 0, 0}, grey);
    s.quad({0, 0, 0}, {0, 0, 1}, {0, 1, 1}, {0, 1, 0}, red);
    s.quad({1, 0, 0}, {1, 1, 0}, {1, 1, 1}, {1, 0, 1}, green);
    s.sphere({0.3f, 0.2f, 0.62f}, 0.2f, grey);
    s.sphere({0.7f, 0.2f, 0.35f}, 0.2f, mirror);
    return rtgpu::Camera{{0.5f, 0.5f, 2.4f}, {0.5f, 0.5f, 0.f}, {0, 1, 0}, (float)(40.0 * M_PI / 180.0), w, h};
}

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s out.raw [w h sx sy mode depth seed]\n", argv[0]);
        return 2;
    }
    if (argc > 3 && std::string(argv[2]) == "--scene") {
        int sx = argc > 4 ? std::atoi(argv[4]) : 2, sy = argc > 5 ? std::atoi(argv[5]) : 2;
        int mode = argc > 6 ? std::atoi(argv[6]) : RT_MODE_UNBIASED, depth = argc > 7 ? std::atoi(argv[7]) : 8;
        unsigned long seed = argc > 8 ? std::strtoul(argv[8], nullptr, 10) : 3;
        try {
            rtgpu::SceneFile file(argv[3]);
            const rt_scene_desc& d = file.scene();
            rtgpu::GpuUniDiPathTracingIntegrator integrator(0.5f, mode, depth);
            integrator.setScene(d);
            rtgpu::PerPixelRenderer renderer(integrator, sx, sy, seed);
            return write_film(argv[1], renderer.render(file.camera()));
        } catch (const rtgpu::RtError& e) {
            std::fprintf(stderr, "error %d: %s\n", e.status, e.what());
            return 1;
        }
    }
    if (argc > 3 && std::string(argv[2]) == "--rays") {
        int mode = argc > 4 ? std::atoi(argv[4]) : RT_MODE_COMPAT, depth = argc > 5 ? std::atoi(argv[5]) : 8;
        uint32_t key = argc > 6 ? (uint32_t)std::strtoul(argv[6], nullptr, 10) : 0;
        unsigned long seed = argc > 7 ? std::strtoul(argv[7], nullptr, 10) : 3;
        std::vector<float> raw;
        if (FILE* f = std::fopen(argv[3], "rb")) {
            float v[8];
            while (std::fread(v, sizeof(float), 8, f) == 8) raw.insert(raw.end(), v, v + 8);
            std::fclose(f);
        }
        try {
            rtgpu::SceneBuilder s;
            box(s, 8, 8);
            rtgpu::GpuUniDiPathTracingIntegrator integ(0.5f, mode, depth);
            integ.setScene(s);
            std::vector<float> out;
            for (size_t i = 0; i < raw.size() / 8; ++i) {
                const float* v = &raw[8 * i];
                rtgpu::Ray r;
                r.o = {v[0], v[1], v[2]};
                r.from = v[3];
                r.d = {v[4], v[5], v[6]};
                r.to = v[7];
                rtgpu::Vec3 c = integ.integrate(r, key + (uint32_t)i, seed);
                out.insert(out.end(), {c.x, c.y, c.z});
            }
            return write_film(argv[1], out);
        } catch (const rtgpu::RtError& e) {
            std::fprintf(stderr, "%s\n", e.what());
            return 1;
        }
    }
    if (argc > 3 && std::string(argv[2]) == "--mlt") {
        const std::string kind = argv[3];
        int w = argc > 4 ? std::atoi(argv[4]) : 64, h = argc > 5 ? std::atoi(argv[5]) : 64;
        int muts = argc > 6 ? std::atoi(argv[6]) : 16;
        unsigned long seed = argc > 7 ? std::strtoul(argv[7], nullptr, 10) : 3;
        try {
            rtgpu::SceneBuilder s;
            rtgpu::Camera cam = box(s, w, h);
            std::vector<float> film;
            if (kind == "metropolis") {
                rtgpu::MetropolisRenderer r(0.1f, 0.5f, 10000, muts, seed);
                r.setScene(s);
                film = r.render(cam);
            } else if (kind == "metropolisflt") {
                rtgpu::MetropolisFltRenderer r(3, 0, 0.1f, 0.5f, 10000, muts, seed);
                r.setScene(s);
                film = r.render(cam);
            } else if (kind == "energyredist") {
                rtgpu::EnergyRedistributionRenderer r(muts, 2, 2, seed);
                r.setScene(s);
                film = r.render(cam);
            } else {
                std::fprintf(stderr, "unknown renderer %s\n", kind.c_str());
                return 2;
            }
            return write_film(argv[1], film);
        } catch (const rtgpu::RtError& e) {
            std::fprintf(stderr, "error %d: %s\n", e.status, e.what());
            return 1;
        }
    }
    int w = argc > 2 ? std::atoi(argv[2]) : 64, h = argc > 3 ? std::atoi(argv[3]) : 64;
    int sx = argc > 4 ? std::atoi(argv[4]) : 2, sy = argc > 5 ? std::atoi(argv[5]) : 2;
    int mode = argc > 6 ? std::atoi(argv[6]) : RT_MODE_UNBIASED, depth = argc > 7 ? std::atoi(argv[7]) : 8;
    unsigned long seed = argc > 8 ? std::strtoul(argv[8], nullptr, 10) : 3;
    using rtgpu::Vec3;
    try {
        rtgpu::SceneBuilder s;
        rtgpu::Camera cam = box(s, w, h);

        rtgpu::GpuUniDiPathTracingIntegrator integrator(0.5f, mode, depth);
        integrator.setScene(s);
        rtgpu::PerPixelRenderer renderer(integrator, sx, sy, seed);
        return write_film(argv[1], renderer.render(cam));
    } catch (const rtgpu::RtError& e) {
        std::fprintf(stderr, "error %d: %s\n", e.status, e.what());
        return 1;
    }
    return 0;
}
